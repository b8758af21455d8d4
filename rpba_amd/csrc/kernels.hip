// HIP kernels (gfx950) of the plain-FP64 robust LM Schur solve.
//
// Layout in HBM (one bundle block):
//   track entries q (point-major, the order of the reference's pointXXv):
//     tcam[q], tpt[q] int32; ox[q], oy[q] float (xg); weights as SoA floats
//     (wx, wy, wxy) for the current Wij and the original Wijoldv; key[q] the
//     float robust key sig0rmv; flag[q] ISflag.
//   points: X[3*i] accepted (XPMatrices point part), Xn trial.
//   per entry 9x3 W = Bc^T W Bp and Y = W V^-1 as 27 doubles in 32-double
//     (256-byte, cache-line aligned) records.
//   cameras: uniform 9-slot parameter blocks (rot a,b,c | C x,y,z | d0, d1, f)
//     with a per-camera mask of the estimated slots (image 0 pose and one C
//     component of the gauge camera are fixed, rbacore.cpp:1042-1069).
//   S: diagonal blocks Sd[k] (81 doubles, row-major) and the upper
//     off-diagonal blocks Su[b] of the covisibility list.
// Every reduction runs in a fixed order: results are bitwise reproducible.
#include <hip/hip_runtime.h>
#include <cstdlib>

#include "kernels.h"

namespace rpba {

// ---------------------------------------------------------------------------
// per-observation math
// ---------------------------------------------------------------------------

struct CamV {
    double R[9], C[3], d0, d1, f;
};

__device__ __forceinline__ void load_cam(const CamPtrs &cp, int k, CamV &c)
{
#pragma unroll
    for (int j = 0; j < 9; ++j) c.R[j] = cp.Rn[9 * k + j];
#pragma unroll
    for (int j = 0; j < 3; ++j) c.C[j] = cp.Cn[3 * k + j];
    c.d0 = cp.In[3 * k + 0];
    c.d1 = cp.In[3 * k + 1];
    c.f = cp.In[3 * k + 2];
}

struct Ev {
    double u, v, xs3, su, sv, d1, d2;
    bool nan;
};

// projection and residual (CalcCorrections rbacore.cpp:1238-1307,
// RadialReconstructionImage rpba.h:850-857)
__device__ __forceinline__ void eval_obs(const CamV &c, double X0, double X1, double X2, float ox, float oy,
                                         double defdiff, Ev &e)
{
    const double p0 = X0 + c.C[0], p1 = X1 + c.C[1], p2 = X2 + c.C[2];
    double ex0 = c.R[0] * p0 + c.R[1] * p1 + c.R[2] * p2;
    double ex1 = c.R[3] * p0 + c.R[4] * p1 + c.R[5] * p2;
    double xs3 = c.R[6] * p0 + c.R[7] * p1 + c.R[8] * p2;
    if (fabs(xs3) < 1.e-3)
        xs3 = (xs3 < 0.) ? -1.e3 : 1.e3;
    else
        xs3 = 1. / xs3;
    ex0 *= xs3;
    ex1 *= xs3;
    e.u = ex0;
    e.v = ex1;
    e.xs3 = xs3;
    const double dist = ex0 * ex0 + ex1 * ex1;
    const double z = 1. + (c.d0 * dist + c.d1 * dist * dist);
    e.su = ex0 * z;
    e.sv = ex1 * z;
    double d1 = (double)ox - e.su * c.f;
    double d2 = (double)oy - e.sv * c.f;
    if (d1 > defdiff) d1 = defdiff;
    if (d1 < -defdiff) d1 = -defdiff;
    if (d2 > defdiff) d2 = defdiff;
    if (d2 < -defdiff) d2 = -defdiff;
    e.nan = isnan(d1) || isnan(d2) || isinf(d1) || isinf(d2);
    if (e.nan) d1 = d2 = 0.;
    e.d1 = d1;
    e.d2 = d2;
}

// calibrated Jacobian dr/dtheta (CalcDerivatives rbacore.cpp:1420-1616,
// eCalcCalibratedDerivativesX/3 1620-1673) in the uniform 9-slot layout;
// unestimated slots give zero columns.  T = double is the plain form; T =
// float evaluates the same expressions in FP32 from the FP64 projection.
template <typename T>
__device__ __forceinline__ void jac_obs(const CamV &c, const Ev &e, unsigned mask, T Bc[2][9], T Bp[2][3])
{
    const T one = (T)1, zero = (T)0;
    const T u = (T)e.u, v = (T)e.v, xs3 = (T)e.xs3;
    const T x1 = u * xs3, x2 = v * xs3;
    const T xs31 = one / xs3, xs32 = xs31 * xs31;
    const T k2 = (T)c.d0, k4 = (T)c.d1, f = (T)c.f;
    const T xy2 = u * u + v * v, xy4 = xy2 * xy2;
    const T xxpc = k4 * xy4 + u * (2 * k2 * u + 4 * k4 * u * xy2) + k2 * xy2 + one;
    const T xypc = u * (4 * k4 * xy2 * v + 2 * k2 * v);
    const T yxpc = (4 * k4 * u * xy2 + 2 * k2 * u) * v;
    const T yypc = v * (2 * k2 * v + 4 * k4 * xy2 * v) + k4 * xy4 + k2 * xy2 + one;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T c0 = -(T)c.R[0 * 3 + j] * xs3 + (T)c.R[2 * 3 + j] * x1;
        const T c1 = -(T)c.R[1 * 3 + j] * xs3 + (T)c.R[2 * 3 + j] * x2;
        const T xp = xxpc * c0 + xypc * c1;
        const T yp = yypc * c1 + yxpc * c0;
        Bp[0][j] = xp * f + yp * zero;
        Bp[1][j] = yp * f;
    }
    T rc[2][3];
    rc[0][0] = x2 * x1 * xs32;
    rc[1][0] = one + x2 * x2 * xs32;
    rc[0][1] = -one - x1 * x1 * xs32;
    rc[1][1] = -x2 * x1 * xs32;
    rc[0][2] = x2 * xs31;
    rc[1][2] = -x1 * xs31;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T xp = xxpc * rc[0][j] + xypc * rc[1][j];
        const T yp = yypc * rc[1][j] + yxpc * rc[0][j];
        Bc[0][j] = xp * f + yp * zero;
        Bc[1][j] = yp * f;
        Bc[0][3 + j] = Bp[0][j];
        Bc[1][3 + j] = Bp[1][j];
    }
    const T xyp2 = u * u + v * v, xyp4 = xyp2 * xyp2;
    Bc[0][6] = -u * xyp2 * f - v * xyp2 * zero;
    Bc[1][6] = -v * xyp2 * f;
    Bc[0][7] = -u * xyp4 * f - v * xyp4 * zero;
    Bc[1][7] = -v * xyp4 * f;
    Bc[0][8] = -(T)e.su;
    Bc[1][8] = -(T)e.sv;
#pragma unroll
    for (int j = 0; j < 9; ++j)
        if (!(mask & (1u << j))) Bc[0][j] = Bc[1][j] = zero;
}

// point group c: {first point, end point, first entry, end entry}, one load
// when the per-group table is there
__device__ __forceinline__ int4 point_group(const int4 *pg4, const int *pgrp, const int *toff, int c)
{
    if (pg4) return pg4[c];
    const int p0 = pgrp[c], p1 = pgrp[c + 1];
    return make_int4(p0, p1, toff[p0], toff[p1]);
}

// 3x3 inverse by cofactors (adjugate / determinant)
__device__ __forceinline__ void inv3(const double m[9], double r[9])
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

// ---------------------------------------------------------------------------
// deterministic reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// block-wide sum, fixed order; every thread gets the result.  sh holds
// blockDim/64 doubles.
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.;
    for (int i = 0; i < nw; ++i) s += sh[i];
    return s;
}

// ---------------------------------------------------------------------------
// linearisation: points (V, g_p, W, Y), one thread per point
// ---------------------------------------------------------------------------

// calibrated point columns only (the Bp part of jac_obs)
template <typename T>
__device__ __forceinline__ void jac_point(const CamV &c, const Ev &e, T Bp[2][3])
{
    const T one = (T)1, zero = (T)0;
    const T u = (T)e.u, v = (T)e.v, xs3 = (T)e.xs3;
    const T x1 = u * xs3, x2 = v * xs3;
    const T k2 = (T)c.d0, k4 = (T)c.d1, f = (T)c.f;
    const T xy2 = u * u + v * v, xy4 = xy2 * xy2;
    const T xxpc = k4 * xy4 + u * (2 * k2 * u + 4 * k4 * u * xy2) + k2 * xy2 + one;
    const T xypc = u * (4 * k4 * xy2 * v + 2 * k2 * v);
    const T yxpc = (4 * k4 * u * xy2 + 2 * k2 * u) * v;
    const T yypc = v * (2 * k2 * v + 4 * k4 * xy2 * v) + k4 * xy4 + k2 * xy2 + one;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T c0 = -(T)c.R[0 * 3 + j] * xs3 + (T)c.R[2 * 3 + j] * x1;
        const T c1 = -(T)c.R[1 * 3 + j] * xs3 + (T)c.R[2 * 3 + j] * x2;
        const T xp = xxpc * c0 + xypc * c1;
        const T yp = yypc * c1 + yxpc * c0;
        Bp[0][j] = xp * f + yp * zero;
        Bp[1][j] = yp * f;
    }
}

// orders LDS accesses of one wavefront (lanes exchange values through LDS)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 16-byte piece j of a U-value record (zero padding after value U - 1)
template <int U>
__device__ __forceinline__ uint4 rec_piece(const double *v, int j)
{
    double2 d;
    d.x = (2 * j < U) ? v[2 * j] : 0.;
    d.y = (2 * j + 1 < U) ? v[2 * j + 1] : 0.;
    return __builtin_bit_cast(uint4, d);
}
template <int U>
__device__ __forceinline__ uint4 rec_piece(const float *v, int j)
{
    float4 f;
    f.x = (4 * j < U) ? v[4 * j] : 0.f;
    f.y = (4 * j + 1 < U) ? v[4 * j + 1] : 0.f;
    f.z = (4 * j + 2 < U) ? v[4 * j + 2] : 0.f;
    f.w = (4 * j + 3 < U) ? v[4 * j + 3] : 0.f;
    return __builtin_bit_cast(uint4, f);
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int kMfChunkWaves = 4;  // chunks (wavefronts) per workgroup of k_mf_chunks

// The records q0 .. q0 + 63 of a wavefront as one contiguous span of 16-byte
// pieces: load instruction i, lane l takes piece 64 i + l of the span (one
// contiguous kilobyte per instruction); pieces past record qlast (a valid
// record) are clamped onto its last piece and dropped.  span_issue issues every
// load (one round trip; the caller may issue loads of its own behind them),
// span_transpose hands each record to its lane through the wavefront's LDS:
// the span's image, RP records (RP NP pieces) per pass, lanes of the pass's
// records read theirs.  Every lane of the wavefront calls both; lanes past nq
// get nothing.
// E: elements stored per record (kRec for Y and W V^-1 style records, kWRec
// for the factored W records); U: the values used (the rest is padding)
template <typename R, int E = kRec>
struct RecSpan {
    static constexpr int U = E == kRec ? 27 : E;
    static constexpr int NP = E * (int)sizeof(R) / 16;     // pieces per record: 14 / 12 (FP64) or 7 / 6 (FP32)
    static constexpr int NPASS = sizeof(R) == 8 ? 2 : 1;
    static constexpr int RP = 64 / NPASS;                  // records per pass
    static constexpr int IPP = NP / NPASS;                 // load instructions per pass (RP NP / 64)
    static constexpr int LDS = RP * NP;                    // pieces of LDS per wavefront
    u32x4 raw[NP];                                         // 64 records x NP pieces / 64 lanes
};
static_assert(kRec * 4 % 16 == 0 && (kRec * 8 / 16) % 2 == 0, "records are whole pieces, two passes of whole loads");
static_assert(kWRec * 4 % 16 == 0 && (kWRec * 8 / 16) % 2 == 0, "records are whole pieces, two passes of whole loads");

template <typename R, int E = kRec>
__device__ __forceinline__ void span_issue(const R *recs, int q0, int qlast, RecSpan<R, E> &w, int lane)
{
    using S = RecSpan<R, E>;
    // piece 64 i + lane of the span, clamped onto the last piece of record
    // qlast (the span's pieces past it are dropped)
    const u32x4 *base = (const u32x4 *)recs + (size_t)q0 * S::NP;
    const int last = (qlast - q0 + 1) * S::NP - 1;
#pragma unroll
    for (int i = 0; i < S::NP; ++i) w.raw[i] = base[min(64 * i + lane, last)];
}

template <typename R, int E = kRec>
__device__ __forceinline__ void span_transpose(const RecSpan<R, E> &w, int nq, R *v, u32x4 *stg, int lane)
{
    using S = RecSpan<R, E>;
    constexpr int U = S::U;
#pragma unroll
    for (int h = 0; h < S::NPASS; ++h) {
#pragma unroll
        for (int i = 0; i < S::IPP; ++i) stg[64 * i + lane] = w.raw[S::IPP * h + i];
        wave_lds_sync();
        const int rl = lane - S::RP * h;
        if (rl >= 0 && rl < S::RP && lane < nq) {
#pragma unroll
            for (int pc = 0; pc < S::NP; ++pc) {
                const u32x4 u = stg[S::NP * rl + pc];
                if constexpr (sizeof(R) == 8) {
                    const double2 d = __builtin_bit_cast(double2, u);
                    if (2 * pc < U) v[2 * pc] = d.x;
                    if (2 * pc + 1 < U) v[2 * pc + 1] = d.y;
                } else {
                    const float4 f = __builtin_bit_cast(float4, u);
                    if (4 * pc < U) v[4 * pc] = f.x;
                    if (4 * pc + 1 < U) v[4 * pc + 1] = f.y;
                    if (4 * pc + 2 < U) v[4 * pc + 2] = f.z;
                    if (4 * pc + 3 < U) v[4 * pc + 3] = f.w;
                }
            }
        }
        wave_lds_sync();
    }
}

// The records of a wavefront's entries q0 .. q0 + nq - 1 (nq <= 64), written
// as whole 128-byte lines: each lane's record goes to the wavefront's LDS eight
// 16-byte pieces at a time, then lane l stores piece l % 8 of record
// 8 p + l / 8, so one store instruction writes eight full lines (a lane
// storing its own record would touch 64 lines per instruction).  Record of
// entry q at rank[q] (or q).  Every lane of the wavefront calls this; lanes
// past the entries (lane >= nq) hold no record.
template <typename R, int E = kRec>
__device__ __forceinline__ void store_records_wave(R *base, const int *rank, int q0, int nq, const R *v, uint4 (*stg)[9],
                                                   int lane)
{
    constexpr int NP = E * (int)sizeof(R) / 16;  // pieces per record: 14 / 12 (FP64) or 7 / 6 (FP32)
    constexpr int U = E == kRec ? 27 : E;
#pragma unroll
    for (int h = 0; h < (NP + 7) / 8; ++h) {
        if (lane < nq) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (8 * h + j < NP) stg[lane][j] = rec_piece<U>(v, 8 * h + j);
        }
        wave_lds_sync();
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const int rr = 8 * p + (lane >> 3), pp = lane & 7;
            if (rr < nq && 8 * h + pp < NP) {
                const size_t rec = rank ? (size_t)rank[q0 + rr] : (size_t)(q0 + rr);
                ((uint4 *)(base + rec * E))[8 * h + pp] = stg[rr][pp];
            }
        }
        wave_lds_sync();
    }
}


// W = Bc^T W Bp (9 x 3) of one entry and, with WithY, Y = W V^-1, in R: the
// values of its records.  The point side (k_lin_records) and the camera side
// of the matrix-free form (k_lin_chunks) both form them here from the same
// inputs, so the two agree bit for bit.
template <typename R, bool WithY>
__device__ __forceinline__ void wy_values(const R (&Bc)[2][9], const R (&Bp)[2][3], R W0, R W1, R W2, const R (&Vr)[9],
                                          R (&Wv)[27], R (&Yv)[27])
{
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const R t0 = Bc[0][r] * W0 + Bc[1][r] * W2;
        const R t1 = Bc[0][r] * W2 + Bc[1][r] * W1;
        const R w0 = t0 * Bp[0][0] + t1 * Bp[1][0];
        const R w1 = t0 * Bp[0][1] + t1 * Bp[1][1];
        const R w2 = t0 * Bp[0][2] + t1 * Bp[1][2];
        Wv[r * 3] = w0;
        Wv[r * 3 + 1] = w1;
        Wv[r * 3 + 2] = w2;
        if constexpr (WithY) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                Yv[r * 3 + cc] = w0 * Vr[0 * 3 + cc] + w1 * Vr[1 * 3 + cc] + w2 * Vr[2 * 3 + cc];
        }
    }
}

// W of one entry as its factors (the kWRec-element record): Bt = W Bc as
// pairs (t0_r, t1_r), formed as in wy_values, then the rows of Bp
template <typename R>
__device__ __forceinline__ void w_factors(const R (&Bc)[2][9], const R (&Bp)[2][3], R W0, R W1, R W2, R (&Wf)[kWRec])
{
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        Wf[2 * r] = Bc[0][r] * W0 + Bc[1][r] * W2;
        Wf[2 * r + 1] = Bc[0][r] * W2 + Bc[1][r] * W1;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Wf[18 + c] = Bp[0][c];
        Wf[21 + c] = Bp[1][c];
    }
}

// element (r, c) of W from its factors: the operations of wy_values, so the
// same bits
template <typename R>
__device__ __forceinline__ R w_elem(const R *Wf, int r, int c)
{
    return Wf[2 * r] * Wf[18 + c] + Wf[2 * r + 1] * Wf[21 + c];
}

// V = sum Bp^T W Bp and g_p = sum Bp^T W r per point, then damping and the
// 3x3 inverse (rbacore.cpp:1341-1344).  A workgroup per point group (first
// entry in a kPtGroup-entry window): lane per entry computes its terms
// (to LDS), then each point's lane adds its entries in order.  Sums are FP64.
// WithW (the matrix-free form): each lane also writes its entry's W record,
// as its factors (w_factors from jac_obs, whose Bp is jac_point's to the bit),
// in whole lines; the records' staging shares the LDS of the entry terms.
template <typename R, bool WithW>
__global__ __launch_bounds__(256) void k_lin_points(LinArgs a)
{
    constexpr size_t kSv = sizeof(double) * 256 * 12, kStg = WithW ? sizeof(uint4) * 4 * 64 * 9 : 0;
    __shared__ __align__(16) unsigned char lds[kSv > kStg ? kSv : kStg];
    double(*sv)[12] = (double(*)[12])lds;
    const int t = threadIdx.x, c = blockIdx.x;
    const int4 grp = point_group(a.pg4, a.pgrp, a.toff, c);
    const int p0 = grp.x, p1 = grp.y, e0 = grp.z, e1 = grp.w;
    if (p0 >= p1) return;
    const int np = p1 - p0;
    const int i = p0 + t;
    int pb = 0, pe = 0;
    if (t < np) {
        pb = a.toff[i];
        pe = a.toff[i + 1];
    }
    double V[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0};
    for (int base = e0; base < e1; base += 256) {
        const int q = base + t;
        double tv[12];
        R Wf[kWRec];
        if (q < e1) {
            const int k = a.tcam[q], ip = a.tpt[q];
            CamV cv;
            load_cam(a.cam, k, cv);
            Ev ev;
            eval_obs(cv, a.X[3 * (size_t)ip], a.X[3 * (size_t)ip + 1], a.X[3 * (size_t)ip + 2], a.ox[q], a.oy[q], a.defdiff, ev);
            R Bp[2][3];
            const R W0 = (R)a.wx[q], W1 = (R)a.wy[q], W2 = (R)a.wxy[q];
            if constexpr (WithW) {
                R Bc[2][9];
                jac_obs<R>(cv, ev, a.cam.mask[k], Bc, Bp);
                w_factors<R>(Bc, Bp, W0, W1, W2, Wf);
            } else {
                jac_point<R>(cv, ev, Bp);
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const R t0 = Bp[0][r] * W0 + Bp[1][r] * W2;
                const R t1 = Bp[0][r] * W2 + Bp[1][r] * W1;
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) tv[r * 3 + cc] = (double)(t0 * Bp[0][cc] + t1 * Bp[1][cc]);
                tv[9 + r] = (double)t0 * ev.d1 + (double)t1 * ev.d2;
            }
        }
        if constexpr (WithW) {
            // the wavefront's 64 entries' records, then the LDS is the terms'
            const int lane = t & 63, wv = t >> 6;
            const int q0 = base + 64 * wv, nq = min(64, max(0, e1 - q0));
            store_records_wave<R, kWRec>((R *)a.Wrec, nullptr, q0, nq, Wf,
                                         (uint4(*)[9])(lds + sizeof(uint4) * 64 * 9 * wv), lane);
            __syncthreads();
        }
        if (q < e1) {
#pragma unroll
            for (int j = 0; j < 12; ++j) sv[t][j] = tv[j];
        }
        __syncthreads();
        if (t < np) {
            const int lo = max(pb, base), hi = min(pe, base + 256);
            for (int x = lo; x < hi; ++x) {
#pragma unroll
                for (int j = 0; j < 9; ++j) V[j] += sv[x - base][j];
#pragma unroll
                for (int j = 0; j < 3; ++j) g[j] += sv[x - base][9 + j];
            }
        }
        __syncthreads();
    }
    if (t >= np) return;
    if (a.gcp_of) {
        // the point's position prior (CalcDerivatives3DGCP rbacore.cpp:1677-1729):
        // B = -I, so V += W3 and g += -W3 (centre - X), added after its
        // observations and before the damping
        const int gk = a.gcp_of[i];
        if (gk >= 0) {
            const double *cg = a.gcp_c + 3 * (size_t)gk;
            const float *w = a.gcp_w + 6 * (size_t)gk;
            const double dd[3] = {cg[0] - a.X[3 * (size_t)i], cg[1] - a.X[3 * (size_t)i + 1], cg[2] - a.X[3 * (size_t)i + 2]};
            const double Ws[9] = {(double)w[0], (double)w[3], (double)w[4], (double)w[3], (double)w[1],
                                  (double)w[5], (double)w[4], (double)w[5], (double)w[2]};
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double b0 = -Ws[r * 3], b1 = -Ws[r * 3 + 1], b2 = -Ws[r * 3 + 2];
#pragma unroll
                for (int c = 0; c < 3; ++c) V[r * 3 + c] += -(-Ws[r * 3 + c]);
                g[r] += b0 * dd[0] + b1 * dd[1] + b2 * dd[2];
            }
        }
    }
    V[0] *= a.mult1;
    V[4] *= a.mult1;
    V[8] *= a.mult1;
    double Vi[9];
    inv3(V, Vi);
#pragma unroll
    for (int j = 0; j < 9; ++j) a.Vinv[9 * (size_t)i + j] = Vi[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) a.gp[3 * (size_t)i + j] = g[j];
    if (a.ycm) {
        // the point's line for the camera side: X, V^-1, g_p, one unused
        double2 *L = (double2 *)(a.ptrec + 16 * (size_t)i);
        L[0] = make_double2(a.X[3 * (size_t)i], a.X[3 * (size_t)i + 1]);
        L[1] = make_double2(a.X[3 * (size_t)i + 2], Vi[0]);
        L[2] = make_double2(Vi[1], Vi[2]);
        L[3] = make_double2(Vi[3], Vi[4]);
        L[4] = make_double2(Vi[5], Vi[6]);
        L[5] = make_double2(Vi[7], Vi[8]);
        L[6] = make_double2(g[0], g[1]);
        L[7] = make_double2(g[2], 0.);
    }
}

// One thread per track entry: recomputes the observation's Jacobian (cheaper
// than storing it) and writes W = Bc^T W Bp and Y = W V^-1 into the entry's
// kRec-element records (R = double: 224 bytes; R = float, the FP32 form: 112
// bytes).  Consecutive lanes write consecutive records, so a wavefront fills
// whole cache lines.  WithY = false (the matrix-free form, a.ycm): only W,
// as its factors (kWRec elements, 192 bytes in FP64); the camera side forms
// Y in camera-major order (k_lin_chunks).
template <typename R, bool WithY>
__global__ __launch_bounds__(128) void k_lin_records(LinArgs a)
{
    __shared__ uint4 stg[2][64][9];  // per wavefront: 64 records' staged pieces
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q0 = q - lane, nq = min(64, a.ntr - q0);
    if (nq <= 0) return;  // the whole wavefront is past the entries
    const int qc = min(q, a.ntr - 1);  // lanes past the entries compute entry ntr - 1 and store nothing
    const int i = a.tpt[qc], k = a.tcam[qc];
    const double X0 = a.X[3 * i], X1 = a.X[3 * i + 1], X2 = a.X[3 * i + 2];
    CamV c;
    load_cam(a.cam, k, c);
    Ev ev;
    eval_obs(c, X0, X1, X2, a.ox[qc], a.oy[qc], a.defdiff, ev);
    R Bc[2][9], Bp[2][3];
    jac_obs<R>(c, ev, a.cam.mask[k], Bc, Bp);
    const R W0 = (R)a.wx[qc], W1 = (R)a.wy[qc], W2 = (R)a.wxy[qc];
    R Vr[9], Wv[27], Yv[27];
    if constexpr (!WithY) {
        R Wf[kWRec];
        w_factors<R>(Bc, Bp, W0, W1, W2, Wf);
        store_records_wave<R, kWRec>((R *)a.Wrec, nullptr, q0, nq, Wf, stg[w], lane);
        return;
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) Vr[j] = (R)a.Vinv[9 * (size_t)i + j];
    wy_values<R, true>(Bc, Bp, W0, W1, W2, Vr, Wv, Yv);
    // whole records, padding included: every cache line is fully written
    store_records_wave<R>((R *)a.Wrec, nullptr, q0, nq, Wv, stg[w], lane);
    store_records_wave<R>((R *)a.Yrec, a.yrank, q0, nq, Yv, stg[w], lane);
}

// ---------------------------------------------------------------------------
// diagonal Schur term per camera, atomic-free: the camera-major entry list is
// cut into chunks of at most 64 entries of one image; one wavefront per chunk
// forms D = sum Y W^T (9x9) and e = sum Y g_p (9) as one matrix product on
// the matrix cores, [Y_1 .. Y_n] (9 x 3n) times [W_1 .. W_n | g_1 .. g_n]^T
// (3n x 10: W^T in columns 0..8, g_p in column 9), as v_mfma_f64_16x16x4f64
// steps over K = 3n (rows and columns 10..15 are zero padding).  Operand
// layout of that instruction: A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4; K index kk = 3 e + c:
// entry e of the chunk, component c of the point.  FP32 records are widened
// to FP64 (products of two floats are exact).  k_lin_cameras adds the chunks
// of an image in order.
// ---------------------------------------------------------------------------

typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kSchurU = 8;  // MFMA steps per load round (32 K values, 10.7 entries or pairs)

template <typename R>
__global__ __launch_bounds__(64) void k_schur_diag(LinArgs a)
{
    const int ch = blockIdx.x;  // chunk: entries [ch_start[ch], ch_start[ch+1]) of one image
    const int lane = threadIdx.x;
    const int i16 = lane & 15, q = lane >> 4;
    const R *Yr = (const R *)a.Yrec, *Wr = (const R *)a.Wrec;
    const int b = a.ch_start[ch], n = a.ch_start[ch + 1] - b;
    __shared__ int sq[64];
    __shared__ double sg[64 * 3];
    if (lane < n) {
        const int qq = a.camq[b + lane];
        const int p = a.cm_pt ? a.cm_pt[b + lane] : a.tpt[qq];
        sq[lane] = qq;
        sg[3 * lane] = a.gp[3 * (size_t)p];
        sg[3 * lane + 1] = a.gp[3 * (size_t)p + 1];
        sg[3 * lane + 2] = a.gp[3 * (size_t)p + 2];
    }
    __syncthreads();
    f64x4 acc = {0., 0., 0., 0.};
    const int kend = 3 * n;
    for (int k0 = 0; k0 < kend; k0 += 4 * kSchurU) {
        double av[kSchurU], bv[kSchurU];
#pragma unroll
        for (int u = 0; u < kSchurU; ++u) {
            const int kk = k0 + 4 * u + q;
            double x = 0., y = 0.;
            if (kk < kend) {
                const int ent = kk / 3, cc = kk - 3 * ent;
                const size_t rec = (size_t)sq[ent] * kRec + 3 * i16 + cc;
                const size_t yrec = a.yrank ? (size_t)(b + ent) * kRec + 3 * i16 + cc : rec;
                if (i16 < 9) {
                    x = (double)Yr[yrec];
                    y = (double)Wr[rec];
                } else if (i16 == 9) {
                    y = sg[3 * ent + cc];
                }
            }
            av[u] = x;
            bv[u] = y;
        }
#pragma unroll
        for (int u = 0; u < kSchurU; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
    }
    // D[i][j] in lane j + 16 (i % 4), element i / 4
    double *D = a.Dd + (size_t)ch * 90;
    if (i16 < 10) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = q + 4 * v;
            if (i < 9) D[(i16 < 9) ? i * 9 + i16 : 81 + i] = acc[v];
        }
    }
}

// Matrix-free form, camera side of the records: one wavefront per chunk (at
// most 64 entries of one image).  Lane e takes camera-major entry b + e: its
// point, observation and weights from the camera-major entry table, its
// point's 128-byte line (X, V^-1, g_p) and the image's parameters.  It
// evaluates the observation and its Jacobian again and forms W with the point
// side's operations (w_factors: W equals the point-major record bit for bit)
// and Y = W V^-1 as Bt^T (Bp V^-1), and Y goes to the entry's camera-major
// record (its factors, kWRec elements) in whole lines.  D = sum Y W^T and
// e = sum Y g_p of the chunk then follow on the matrix cores in the K order of
// k_schur_diag, with the operands staged in LDS 32 entries at a time.  Reads
// per entry: the table (28 bytes) and one point line (128); writes: Y.
template <typename R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_lin_chunks(LinArgs a)
{
    constexpr int kOp = 57;                   // per staged entry: Y (27), W (27), g_p (3)
    __shared__ __align__(16) double opnd[32 * kOp];
    uint4 (*stg)[9] = (uint4 (*)[9])opnd;     // the record stores' staging (64 x 9 pieces) shares it
    static_assert(sizeof(opnd) >= 64 * 9 * sizeof(uint4), "staging fits");
    const int ch = blockIdx.x;
    const int lane = threadIdx.x;
    const int i16 = lane & 15, q = lane >> 4;
    const int4 c4 = a.ch4[ch];  // {first entry, entries, image, first chunk of the image}
    const int b = c4.x, n = c4.y, k = c4.z;
    const int j = b + min(lane, n - 1);  // lanes past the chunk's entries take its last entry and store nothing
    const int p = a.cm_pt[j];
    const float4 wv = a.cm_w[j];
    const double2 *L = (const double2 *)(a.ptrec + 16 * (size_t)p);
    double ln[16];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        const double2 d = L[h];
        ln[2 * h] = d.x;
        ln[2 * h + 1] = d.y;
    }
    CamV c;
    load_cam(a.cam, k, c);
    Ev ev;
    eval_obs(c, ln[0], ln[1], ln[2], a.cm_ox[j], a.cm_oy[j], a.defdiff, ev);
    R Bc[2][9], Bp[2][3];
    jac_obs<R>(c, ev, a.cam.mask[k], Bc, Bp);
    // W as its factors (Bt pairs, Bp: the point-major record's values) and
    // Y = W V^-1 as Bt^T P with P = Bp V^-1 (2 x 3): the Y record holds the
    // Bt pairs and P (kWRec elements), and D below is formed from these same
    // Y values, so the product's own-camera terms cancel D_k exactly
    R Wf[kWRec], Yf[kWRec], Wv[27], Yv[27];
    w_factors<R>(Bc, Bp, (R)wv.x, (R)wv.y, (R)wv.z, Wf);
#pragma unroll
    for (int x = 0; x < 18; ++x) Yf[x] = Wf[x];
#pragma unroll
    for (int sr = 0; sr < 2; ++sr)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            Yf[18 + 3 * sr + cc] = Bp[sr][0] * (R)ln[3 + cc] + Bp[sr][1] * (R)ln[6 + cc] + Bp[sr][2] * (R)ln[9 + cc];
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            Wv[3 * r + cc] = w_elem(Wf, r, cc);
            Yv[3 * r + cc] = w_elem(Yf, r, cc);
        }
    store_records_wave<R, kWRec>((R *)a.Yrec, nullptr, b, n, Yf, stg, lane);
    f64x4 acc = {0., 0., 0., 0.};
    for (int base = 0; base < n; base += 32) {
        const int nr = min(32, n - base);
        if (lane >= base && lane < base + nr) {
            double *o = opnd + kOp * (lane - base);
#pragma unroll
            for (int x = 0; x < 27; ++x) {
                o[x] = (double)Yv[x];
                o[27 + x] = (double)Wv[x];
            }
            o[54] = ln[12];
            o[55] = ln[13];
            o[56] = ln[14];
        }
        wave_lds_sync();
        const int kend = 3 * nr;
        for (int k0 = 0; k0 < kend; k0 += 4 * kSchurU) {
            double av[kSchurU], bv[kSchurU];
#pragma unroll
            for (int u = 0; u < kSchurU; ++u) {
                const int kk = k0 + 4 * u + q;
                double x = 0., y = 0.;
                if (kk < kend) {
                    const int ent = kk / 3, cc = kk - 3 * ent;
                    const double *o = opnd + kOp * ent;
                    if (i16 < 9) {
                        x = o[3 * i16 + cc];
                        y = o[27 + 3 * i16 + cc];
                    } else if (i16 == 9) {
                        y = o[54 + cc];
                    }
                }
                av[u] = x;
                bv[u] = y;
            }
#pragma unroll
            for (int u = 0; u < kSchurU; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
        }
        wave_lds_sync();
    }
    // D[i][j] in lane j + 16 (i % 4), element i / 4
    double *D = a.Dd + (size_t)ch * 90;
    if (i16 < 10) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = q + 4 * v;
            if (i < 9) D[(i16 < 9) ? i * 9 + i16 : 81 + i] = acc[v];
        }
    }
}

// ---------------------------------------------------------------------------
// linearisation: cameras.  One workgroup (256) per camera accumulates
// U = sum Bc^T W Bc (upper 45) and g_c = sum Bc^T W r, then forms
// S_kk = U - D with the diagonal damping (x(1+lambda), +1e-2,
// rbacore.cpp:1164-1170) and b_k = g_c - e; fixed slots become an identity
// row/column with zero right-hand side.
// ---------------------------------------------------------------------------

template <typename R>
__global__ __launch_bounds__(256) void k_lin_cameras(LinArgs a)
{
    const int k = blockIdx.x;
    const int t = threadIdx.x;
    __shared__ double red[4][54];
    __shared__ double Ssh[81];
    __shared__ double bsh[9];
    const unsigned mask = a.cam.mask[k];
    CamV c;
    load_cam(a.cam, k, c);
    const int b = a.camoff[k], e = a.camoff[k + 1];

    double acc[54];
#pragma unroll
    for (int j = 0; j < 54; ++j) acc[j] = 0.;
    const bool cm = a.cm_pt != nullptr;
    for (int j = b + t; j < e; j += 256) {
        int p;
        float ox, oy, w0, w1, w2;
        if (cm) {
            p = a.cm_pt[j];
            ox = a.cm_ox[j];
            oy = a.cm_oy[j];
            const float4 wv = a.cm_w[j];
            w0 = wv.x;
            w1 = wv.y;
            w2 = wv.z;
        } else {
            const int q = a.camq[j];
            p = a.tpt[q];
            ox = a.ox[q];
            oy = a.oy[q];
            w0 = a.wx[q];
            w1 = a.wy[q];
            w2 = a.wxy[q];
        }
        Ev ev;
        eval_obs(c, a.X[3 * p], a.X[3 * p + 1], a.X[3 * p + 2], ox, oy, a.defdiff, ev);
        R Bc[2][9], Bp[2][3];
        jac_obs<R>(c, ev, mask, Bc, Bp);
        const R W0 = (R)w0, W1 = (R)w1, W2 = (R)w2;
        R Bt[9][2];
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            Bt[r][0] = Bc[0][r] * W0 + Bc[1][r] * W2;
            Bt[r][1] = Bc[0][r] * W2 + Bc[1][r] * W1;
        }
        int n = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r)
#pragma unroll
            for (int cc = r; cc < 9; ++cc) acc[n++] += (double)(Bt[r][0] * Bc[0][cc] + Bt[r][1] * Bc[1][cc]);
#pragma unroll
        for (int r = 0; r < 9; ++r) acc[45 + r] += (double)Bt[r][0] * ev.d1 + (double)Bt[r][1] * ev.d2;
    }
    const int w = t >> 6, lane = t & 63;
#pragma unroll
    for (int j = 0; j < 54; ++j) {
        const double s = wave_sum(acc[j]);
        if (lane == 0) red[w][j] = s;
    }
    // diagonal Schur terms: the image's chunks in order
    __shared__ double Dsh[90];
    if (t < 90) {
        double sd = 0.;
        for (int ch = a.ch_off[k]; ch < a.ch_off[k + 1]; ++ch) sd += a.Dd[(size_t)ch * 90 + t];
        Dsh[t] = sd;
    }
    __syncthreads();
    const double *D = Dsh;
    if (t < 81) {
        const int r = t / 9, cc = t % 9;
        const int rr = r < cc ? r : cc, c2 = r < cc ? cc : r;
        const int n = rr * 9 - rr * (rr - 1) / 2 + (c2 - rr);  // index in the packed upper triangle
        const double U = red[0][n] + red[1][n] + red[2][n] + red[3][n];
        double s = U - D[t];
        if (r == cc) {
            s *= a.mult1;
            s += 1.e-2;
        }
        const bool fr = (mask >> r) & 1u, fc = (mask >> cc) & 1u;
        if (!fr || !fc) s = (r == cc) ? 1. : 0.;
        Ssh[t] = s;
    } else if (t >= 128 && t < 137) {
        const int r = t - 128;
        const double gc = red[0][45 + r] + red[1][45 + r] + red[2][45 + r] + red[3][45 + r];
        double s = gc - D[81 + r];
        if (!((mask >> r) & 1u)) s = 0.;
        bsh[r] = s;
    }
    __syncthreads();
    if (t < 81) a.Sd[81 * (size_t)k + t] = Ssh[t];
    if (t < 9) a.bvec[9 * (size_t)k + t] = bsh[t];
    if (a.Dk && t < 81) a.Dk[81 * (size_t)k + t] = D[t];
}

// ---------------------------------------------------------------------------

// Off-diagonal blocks, atomic-free, one wavefront per strip slab: lane 9 g + i
// accumulates row i of the slab's block g, S_kl[i][j] = -sum over the block's
// pairs p (in order) and c = 0..2 of Y_a[i][c] W_b[j][c], as FP64 fused
// multiply-adds in that order (the K order of a matrix-core product of
// [Y_a1 .. Y_an] and [W_b1 .. W_bn]^T, which measured slower here: a block has
// 2 to 30 pairs, far too few to fill a 16 x 16 x 4 tile step).  Y_a's row i is the lane's own 24 bytes of
// the record; W_b goes through the wavefront's LDS (each lane of the block
// stores three of its 27 values).  The next pair's index and records are in
// flight while the current pair is added; every load is issued (indices
// clamped into the arrays) so the compiler counts them exactly.  Stores:
// for each column j the slab's rows are one run of 9 m doubles.  Slabs go in
// contiguous ranges per XCD, four wavefronts per workgroup, grid-stride.
template <typename R>
__global__ __launch_bounds__(256) void k_schur_slabs(SchurArgs a)
{
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int g = lane / 9, i = lane - 9 * g;  // g = 7: lane 63, idle
    __shared__ double Wsh[4][kSlab + 1][27];
    const R *Yr = (const R *)a.Yrec, *Wr = (const R *)a.Wrec;
    const int per = (a.nslab + 7) / 8;
    const int x = (int)(blockIdx.x % 8);
    const int stride = (int)(gridDim.x / 8) * 4;
    for (int it = (int)(blockIdx.x / 8) * 4 + w; it < per; it += stride) {
        const int s = x * per + it;
        if (s >= a.nslab) break;
        const int base = a.slab_bs[s], m = a.slab_bs[s + 1] - base;
        int p0 = 0, np = 0;
        if (g < m) {
            const int b = a.sperm[base + g];
            p0 = a.poff[b];
            np = a.poff[b + 1] - p0;
        }
        // pair p's index (clamped into the block's list, or pair 0 for idle lanes)
        auto pidx = [&](int p) -> long long {
            return np > 0 ? (long long)p0 + min(p, np - 1) : 0;
        };
        double acc[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) acc[j] = 0.;
        unsigned long long ab = a.pab[pidx(0)];
        unsigned long long abn = a.pab[pidx(1)];
        R y[3], wv[3], yn[3], wn[3];
        {
            const R *yr = Yr + (size_t)(ab >> 32) * kRec + 3 * i;
            const R *wr = Wr + (size_t)(ab & 0xffffffffull) * kRec + 3 * i;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                y[c] = yr[c];
                wv[c] = wr[c];
            }
        }
        for (int p = 0;; ++p) {
            const bool act = g < m && p < np;
            if (!__any(act)) break;
            // in flight: pair p + 2's index, pair p + 1's records
            const unsigned long long ab2 = a.pab[pidx(p + 2)];
            {
                const R *yr = Yr + (size_t)(abn >> 32) * kRec + 3 * i;
                const R *wr = Wr + (size_t)(abn & 0xffffffffull) * kRec + 3 * i;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    yn[c] = yr[c];
                    wn[c] = wr[c];
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) Wsh[w][g][3 * i + c] = (double)wv[c];
            wave_lds_sync();
            if (act) {
#pragma unroll
                for (int j = 0; j < 9; ++j)
#pragma unroll
                    for (int c = 0; c < 3; ++c) acc[j] = fma((double)y[c], Wsh[w][g][3 * j + c], acc[j]);
            }
            wave_lds_sync();
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                y[c] = yn[c];
                wv[c] = wn[c];
            }
            abn = ab2;
        }
        if (g < m) {
#pragma unroll
            for (int j = 0; j < 9; ++j) a.Su[81 * (size_t)base + (size_t)(9 * m) * j + 9 * g + i] = -acc[j];
        }
    }
}

// ---------------------------------------------------------------------------
// PCG (SparsePCG rbacore.cpp:1807-1914) with device-held scalars
// ---------------------------------------------------------------------------

// block-Jacobi inverse by LU with partial pivoting, one thread per camera
__global__ __launch_bounds__(64) void k_bj_inverse(int nc, const double *Sd, double *Cinv)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    double A[81];
    int perm[9];
    for (int j = 0; j < 81; ++j) A[j] = Sd[81 * (size_t)k + j];
    for (int j = 0; j < 9; ++j) perm[j] = j;
    for (int kk = 0; kk < 9; ++kk) {
        int p = kk;
        double best = fabs(A[kk * 9 + kk]);
        for (int i = kk + 1; i < 9; ++i)
            if (fabs(A[i * 9 + kk]) > best) { best = fabs(A[i * 9 + kk]); p = i; }
        if (p != kk) {
            for (int j = 0; j < 9; ++j) { const double tt = A[kk * 9 + j]; A[kk * 9 + j] = A[p * 9 + j]; A[p * 9 + j] = tt; }
            const int tt = perm[kk]; perm[kk] = perm[p]; perm[p] = tt;
        }
        const double piv = A[kk * 9 + kk];
        if (piv != 0.)
            for (int i = kk + 1; i < 9; ++i) {
                const double l = A[i * 9 + kk] / piv;
                A[i * 9 + kk] = l;
                for (int j = kk + 1; j < 9; ++j) A[i * 9 + j] -= l * A[kk * 9 + j];
            }
    }
    for (int col = 0; col < 9; ++col) {
        double y[9];
        for (int i = 0; i < 9; ++i) {
            double s = (perm[i] == col) ? 1. : 0.;
            for (int j = 0; j < i; ++j) s -= A[i * 9 + j] * y[j];
            y[i] = s;
        }
        for (int i = 8; i >= 0; --i) {
            double s = y[i];
            for (int j = i + 1; j < 9; ++j) s -= A[i * 9 + j] * y[j];
            y[i] = s / A[i * 9 + i];
        }
        for (int i = 0; i < 9; ++i) Cinv[81 * (size_t)k + i * 9 + col] = y[i];
    }
}

// h = C r for camera k (thread-local 9x9 mat-vec)
__device__ __forceinline__ void apply_cinv(const double *Cinv, const double *r, double *h, int k)
{
    const double *C = Cinv + 81 * (size_t)k;
    double rr[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) rr[j] = r[9 * (size_t)k + j];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double s = 0.;
#pragma unroll
        for (int j = 0; j < 9; ++j) s += C[i * 9 + j] * rr[j];
        h[9 * (size_t)k + i] = s;
    }
}

// x = 0, r = b, h = C b, d = h, best = b (exivold starts as a copy of b),
// rho = r.h.  Single workgroup of 1024.  Scalars live in S[0] (launch 0).
// PCG set-up, thread per camera: x = 0, r = best = b, h = d = C r, and the
// workgroup's share of r.h (the camera's nine terms in order, then the
// workgroup's cameras in order) into part[blockIdx.x]
__global__ __launch_bounds__(256) void k_pcg_init(PcgArgs a)
{
    __shared__ double sh[4];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.;
    if (k < a.nc) {
        const size_t j0 = 9 * (size_t)k;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const double bj = a.b[j0 + j];
            a.x[j0 + j] = 0.;
            a.r[j0 + j] = bj;
            a.best[j0 + j] = bj;
        }
        apply_cinv(a.Cinv, a.r, a.h, k);
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const double hj = a.h[j0 + j];
            a.d[j0 + j] = hj;
            s += a.r[j0 + j] * hj;
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) a.part[blockIdx.x] = s;
}

// rho = the shares in workgroup order; the scalars of both slots
__global__ __launch_bounds__(1024) void k_pcg_init_fin(PcgArgs a, int nb)
{
    __shared__ double sh[16];
    double s = 0.;
    const int per = (nb + (int)blockDim.x - 1) / (int)blockDim.x;
    for (int i = threadIdx.x * per; i < min(nb, ((int)threadIdx.x + 1) * per); ++i) s += a.part[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        PcgScalars z;
        z.rho = s;
        z.m0 = 1.e50;
        z.mold = 1.e50;
        z.k = 0;
        z.done = 0;
        z.iters = 0;
        z.nan_flag = 0;
        a.S[0] = z;
        a.S[1] = z;
    }
}

// S d reading every stored block once, from the strip slabs (SuMap).
// k_spmv_slabs: persistent, and every wavefront runs on its own (no workgroup
// barrier inside the chunk loop).  Wavefront w of the grid takes chunks w,
// w + W, ... (W = all wavefronts of the launch; launch j takes them in reverse
// when j is odd, so the blocks read last by one launch, still held by the
// memory-side cache when S is not much larger than it, are read first by the
// next).  A chunk is up to slab_chunk consecutive slabs of one strip (rows
// k0 .. k0 + 6).  Lane 9 g + r holds row r of the block (k0 + g, l) of each
// slab: nine loads, one per column, each a run of 9 m consecutive doubles over
// the slab's m blocks (lanes of absent rows load nothing and count as 0).
// (S_kl d_l)_r is added over the chunk's slabs in the lane's register and
// stored once per chunk (rowpart).  For S_kl^T d_k the products
// S_kl[r][j] d_k[r] go through the wavefront's LDS, where lane 9 g + j adds
// column j over r in order; the slab's blocks' shares are then added over g
// in order into one column share per slab (colpart).  After its chunks a
// workgroup forms the diagonal products yD_k = S_kk d_k of cameras
// k = blockIdx.x, blockIdx.x + gridDim.x, ..., and writes its share of
// d.S d = sum_k d_k.yD_k + 2 sum_(k<l) d_k.S_kl d_l.
// k_pcg_a (below) gathers row k of S d from yD, the row partials of its
// strip's chunks and the column shares of the slabs of column k.
constexpr int kSpmvWaves = 4;  // wavefronts per workgroup (256 threads)


// The read-only arrays of the slab walk (and d) are separate restrict kernel
// arguments: memory the kernel provably never writes, so their wave-uniform
// loads go through the scalar unit and never wait on the vector loads of S.

// A wavefront's walk over its slabs: chunk ci of its sequence (ci = w, w + W,
// ...), chunk id c of strip st, slab q, up to e.  ok = 0 past the last slab;
// q is then slab 0, so that the loads of that position stay in bounds.
// With a.strip_live set, the chunks of strips at 0 are passed over (CI moves
// on to the next chunk of the sequence that is live).
#define SPMV_CHUNK_START(CI, C, Q, E, ST, OK)                                         \
    {                                                                                 \
        for (;;) {                                                                    \
            const int ci_ = (CI);                                                     \
            OK = ci_ < a.nchunk;                                                      \
            C = 0; Q = 0; E = 0; ST = 0;                                              \
            if (!OK) break;                                                           \
            C = rev ? a.nchunk - 1 - ci_ : ci_;                                       \
            ST = chunk_strip[C];                                                      \
            if (a.strip_live && a.strip_live[ST] == 0) {                              \
                CI = ci_ + W;                                                         \
                continue;                                                             \
            }                                                                         \
            Q = strip_s0[ST] + a.slab_chunk * (C - chunk_off[ST]);                    \
            E = min(Q + a.slab_chunk, strip_s0[ST + 1]);                              \
            break;                                                                    \
        }                                                                             \
    }
// slab Q's first slot B, block count M, rows present K and column L, loaded
// (scalar) when the position is found, a slab ahead of their use
#define SPMV_META(Q, B, M, K, L)                                                      \
    {                                                                                 \
        B = slab_bs[Q];                                                               \
        M = slab_bs[(Q) + 1] - B;                                                     \
        K = slab_mask[Q];                                                             \
        L = slab_l[Q];                                                                \
    }
// position after (CI, C, Q, E, ST, OK) into (NCI, NC, NQ, NE, NST, NOK), and
// the new slab's metadata
#define SPMV_ADVANCE(CI, C, Q, E, ST, OK, NCI, NC, NQ, NE, NST, NOK, NB, NM, NK, NL)  \
    {                                                                                 \
        if (OK && Q + 1 < E) {                                                        \
            NCI = CI; NC = C; NQ = Q + 1; NE = E; NST = ST; NOK = 1;                  \
        } else if (OK) {                                                              \
            NCI = CI + W;                                                             \
            SPMV_CHUNK_START(NCI, NC, NQ, NE, NST, NOK);                              \
        } else {                                                                      \
            NCI = CI; NC = C; NQ = Q; NE = E; NST = ST; NOK = 0;                      \
        }                                                                             \
        SPMV_META(NQ, NB, NM, NK, NL);                                                \
    }
// the nine column entries of the lane's row in slab Q and d_k[r] of that row;
// every lane loads (absent rows read the slab's first block and are zeroed
// when used), so the compiler counts the loads exactly
#define SPMV_LOAD(B, M, K, ST, V, DK)                                                 \
    {                                                                                 \
        const int base_ = (B), m_ = (M), mask_ = (K);                                 \
        const bool pr_ = g < kSlab && ((mask_ >> g) & 1);                             \
        const int rank_ = pr_ ? __popc(mask_ & ((1 << g) - 1)) : 0;                   \
        const double *src_ = a.Su + 81 * (size_t)base_ + 9 * rank_ + r;               \
        _Pragma("unroll") for (int cc = 0; cc < 9; ++cc) V[cc] = src_[(size_t)(9 * m_) * cc]; \
        const int k_ = min(kSlab * (ST) + min(g, kSlab - 1), a.nc - 1);               \
        DK = d[9 * (size_t)k_ + r];                                                   \
    }
// one slab: products, column share, and the row partial at the chunk's end
#define SPMV_SLAB(C, Q, E, K, L, V, DK)                                               \
    {                                                                                 \
        const int mask_ = (K), l_ = (L);                                              \
        const bool pr_ = g < kSlab && ((mask_ >> g) & 1);                             \
        double cv_[9];                                                                \
        _Pragma("unroll") for (int cc = 0; cc < 9; ++cc) cv_[cc] = pr_ ? V[cc] : 0.;  \
        const double *dl_ = d + 9 * (size_t)l_;                                       \
        double yv_ = 0.;                                                              \
        _Pragma("unroll") for (int cc = 0; cc < 9; ++cc) yv_ += cv_[cc] * dl_[cc];    \
        yrow += yv_;                                                                  \
        du += DK * yv_;                                                               \
        _Pragma("unroll") for (int jj = 0; jj < 9; ++jj) P[w][g][9 * jj + r] = cv_[jj] * DK; \
        wave_lds_sync();                                                              \
        double tv_ = 0.;                                                              \
        _Pragma("unroll") for (int cc = 0; cc < 9; ++cc) tv_ += P[w][g][9 * r + cc];  \
        T[w][g][r] = tv_;                                                             \
        wave_lds_sync();                                                              \
        if (lane < 9) {                                                               \
            double cs_ = 0.;                                                          \
            _Pragma("unroll") for (int h = 0; h < kSlab; ++h) cs_ += T[w][h][lane];   \
            a.colpart[9 * (size_t)(Q) + lane] = cs_;                                  \
        }                                                                             \
        wave_lds_sync();                                                              \
        if ((Q) + 1 == (E)) {                                                         \
            if (lane < 9 * kSlab) a.rowpart[63 * (size_t)(C) + lane] = yrow;          \
            yrow = 0.;                                                                \
        }                                                                             \
    }

__device__ __forceinline__ void spmv_slabs_body(const PcgArgs &a, int j, const int *__restrict__ chunk_strip,
                                                const int *__restrict__ strip_s0, const int *__restrict__ chunk_off,
                                                const int *__restrict__ slab_bs, const int *__restrict__ slab_mask,
                                                const int *__restrict__ slab_l, const double *__restrict__ d)
{
    const bool rev = (j & 1) != 0;
    const int t = threadIdx.x, lane = t & 63;
    // wavefront index, uniform (held in a scalar register, so the walk over
    // chunks and slabs is scalar)
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int g = lane / 9, r = lane - 9 * g;  // g = 7: lane 63, never present
    const int W = gridDim.x * kSpmvWaves;
    __shared__ double P[kSpmvWaves][kSlab + 1][81];  // S_kl[rr][jj] d_k[rr] at [g][9 jj + rr]
    __shared__ double T[kSpmvWaves][kSlab + 1][9];   // block g's column share
    double du = 0.;    // this lane's share of sum over upper blocks of d_k.S_kl d_l
    double yrow = 0.;  // row k0 + g of the current chunk: sum over its slabs of (S_kl d_l)_r
    // two slabs in flight: register sets A and B alternate
    int ci0, c0, q0, e0, st0, ok0, ci1, c1, q1, e1, st1, ok1;
    int b0, m0, k0, l0, b1, m1, k1, l1;  // the two slabs' metadata
    ci0 = blockIdx.x * kSpmvWaves + w;
    SPMV_CHUNK_START(ci0, c0, q0, e0, st0, ok0);
    SPMV_META(q0, b0, m0, k0, l0);
    SPMV_ADVANCE(ci0, c0, q0, e0, st0, ok0, ci1, c1, q1, e1, st1, ok1, b1, m1, k1, l1);
    double vA[9], vB[9], dkA, dkB;
    // One iteration: slab p0 from set A, slab p1 (if any) from set B, and
    // both sets reloaded unconditionally, so that every path back to the loop
    // head has issued A's loads before B's and the compiler waits only for
    // the set in use.
    if (ok0) {
        SPMV_LOAD(b0, m0, k0, st0, vA, dkA);
        SPMV_LOAD(b1, m1, k1, st1, vB, dkB);
        do {
            int ci2, c2, q2, e2, st2, ok2, ci3, c3, q3, e3, st3, ok3;
            int b2, m2, k2, l2, b3, m3, k3, l3;
            SPMV_ADVANCE(ci1, c1, q1, e1, st1, ok1, ci2, c2, q2, e2, st2, ok2, b2, m2, k2, l2);
            SPMV_ADVANCE(ci2, c2, q2, e2, st2, ok2, ci3, c3, q3, e3, st3, ok3, b3, m3, k3, l3);
            SPMV_SLAB(c0, q0, e0, k0, l0, vA, dkA);
            SPMV_LOAD(b2, m2, k2, st2, vA, dkA);
            if (ok1) SPMV_SLAB(c1, q1, e1, k1, l1, vB, dkB);
            SPMV_LOAD(b3, m3, k3, st3, vB, dkB);
            ci0 = ci2; c0 = c2; q0 = q2; e0 = e2; st0 = st2;
            ci1 = ci3; c1 = c3; q1 = q3; e1 = e3; st1 = st3;
            b0 = b2; m0 = m2; k0 = k2; l0 = l2;
            b1 = b3; m1 = m3; k1 = k3; l1 = l3;
            ok0 = ok1 ? ok2 : 0;
            ok1 = ok3;
        } while (ok0);
    }
    // diagonal blocks
    double dd = 0.;
    for (int k = blockIdx.x; k < a.nc; k += gridDim.x) {
        if (t < 9) {
            const double *B = a.Sd + 81 * (size_t)k + 9 * t;
            const double *dk = a.d + 9 * (size_t)k;
            double y = 0.;
#pragma unroll
            for (int cc = 0; cc < 9; ++cc) y += B[cc] * dk[cc];
            a.Ad[9 * (size_t)k + t] = y;
            dd += dk[t] * y;
        }
    }
    __shared__ double sh[4];
    const double su = block_sum(du, sh);
    const double sd = block_sum(dd, sh);
    if (t == 0) a.dpart[blockIdx.x] = 2. * su + sd;
}

#undef SPMV_CHUNK_START
#undef SPMV_META
#undef SPMV_ADVANCE
#undef SPMV_LOAD
#undef SPMV_SLAB

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_spmv_slabs(PcgArgs a, const int *__restrict__ chunk_strip,
                                                    const int *__restrict__ strip_s0, const int *__restrict__ chunk_off,
                                                    const int *__restrict__ slab_bs, const int *__restrict__ slab_mask,
                                                    const int *__restrict__ slab_l, const double *__restrict__ d)
{
    if (a.S[a.j & 1].done) return;
    spmv_slabs_body(a, a.j, chunk_strip, strip_s0, chunk_off, slab_bs, slab_mask, slab_l, d);
}

// sum of n values in a fixed order; every thread of the block gets it
__device__ __forceinline__ double sum_fixed(const double *v, int n, double *sh)
{
    double s = 0.;
    for (int j = threadIdx.x; j < n; j += blockDim.x) s += v[j];
    return block_sum(s, sh);
}

// Matrix-free S d (no off-diagonal S stored), three launches:
// k_mf_points: t_i = sum over its entries q of W_q^T d_cam(q), stored once per
// point.
// k_mf_chunks: one wavefront per chunk (at most 64 entries of one image):
// lane e gathers t_pt(q) of entry q (through cq_pt) and forms Y_q t_pt(q), the
// wavefront sums the 9 values in lane order into cpart[chunk]; the chunk's
// share of d.S d is
// -d_k.cpart[chunk], plus d_k.(S_kk + D_k) d_k on the image's first chunk,
// whose lanes also store (S_kk + D_k) d_k in Ad[k].  k_pcg_a then forms
// (S d)_k = Ad_k - the image's chunk sums (28 interleaved runs, each in chunk order).
// k_mf_dsum adds the chunk shares of d.S d in a fixed order, in
// kMfDsumGroups contiguous ranges.
template <typename R>
__global__ __launch_bounds__(256) void k_mf_points(PcgArgs a)
{
    // the points of a group (first entry in a kPtGroup-entry window): their
    // entries are taken 256 at a time, one per lane (u_q = W_q^T d_cam(q) to
    // LDS), and each point's lane adds its entries in order and stores t_i
    // (tc).  A lane's image index, its wavefront's records and then the
    // image's d are issued before the first wait.
    if (a.S[a.j & 1].done) return;
    __shared__ double su[256][3];
    const int t = threadIdx.x, c = blockIdx.x;
    const int4 grp = point_group(a.pg4, a.pgrp, a.toff, c);
    const int p0 = grp.x, p1 = grp.y, e0 = grp.z, e1 = grp.w;
    if (p0 >= p1) return;
    const int np = p1 - p0;  // at most kPtGroup / 2 + 1: a point has at least two entries
    int pb = 0, pe = 0;
    if (t < np) {
        pb = a.toff[p0 + t];
        pe = a.toff[p0 + t + 1];
    }
    double t0 = 0., t1 = 0., t2 = 0.;
    const R *Wr = (const R *)a.Wrec;  // W as its factors (kWRec elements)
    __shared__ u32x4 stg[4][RecSpan<R, kWRec>::LDS];  // per wavefront
    const int lane = t & 63, wv = t >> 6;
    for (int base = e0; base < e1; base += 256) {
        const int q = base + t;
        const int cam = a.tcam[min(q, e1 - 1)];
        const int q0 = base + 64 * wv, nq = min(64, max(0, e1 - q0));
        RecSpan<R, kWRec> rw;
        span_issue<R, kWRec>(Wr, q0, e1 - 1, rw, lane);
        double dv[9];
        const double *dp = a.d + 9 * (size_t)cam;
#pragma unroll
        for (int r = 0; r < 9; ++r) dv[r] = dp[r];
        R Wf[kWRec];
        span_transpose<R, kWRec>(rw, nq, Wf, stg[wv], lane);
        if (q < e1) {
            double u0 = 0., u1 = 0., u2 = 0.;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const double dr = dv[r];
                u0 += (double)w_elem(Wf, r, 0) * dr;
                u1 += (double)w_elem(Wf, r, 1) * dr;
                u2 += (double)w_elem(Wf, r, 2) * dr;
            }
            su[t][0] = u0;
            su[t][1] = u1;
            su[t][2] = u2;
        }
        __syncthreads();
        if (t < np) {
            const int lo = max(pb, base), hi = min(pe, base + 256);
            for (int x = lo; x < hi; ++x) {
                t0 += su[x - base][0];
                t1 += su[x - base][1];
                t2 += su[x - base][2];
            }
        }
        __syncthreads();
    }
    // t of the group's points, once per point (k_mf_chunks gathers it for
    // its camera-major entries: 24 bytes per point, which the memory-side
    // cache holds, instead of a copy per entry)
    if (t < np) {
        double *o = a.tc + 3 * (size_t)(p0 + t);
        o[0] = t0;
        o[1] = t1;
        o[2] = t2;
    }
}

// total of the chunk shares into dpart[nch], one workgroup, fixed order
// workgroup g adds the chunk shares of its contiguous range in a fixed
// order; k_pcg_a adds the kMfDsumGroups partials in order
__global__ __launch_bounds__(256) void k_mf_dsum(PcgArgs a)
{
    __shared__ double sh[4];
    if (a.S[a.j & 1].done) return;
    const int per = (a.nch + kMfDsumGroups - 1) / kMfDsumGroups;
    const int c0 = blockIdx.x * per, c1 = min(a.nch, c0 + per);
    double s = 0.;
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) s += a.dpart[c];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) a.dpart[a.nch + blockIdx.x] = s;
}

// the camera-major entry table, from the point-major arrays read in order
// (entry q goes to its camera-major slot camrank[q]): point and observation
// when full, the current weights always; chunk -> its image when full
__global__ void k_cm_table(LinArgs a, CmTable t, int full)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < a.ntr) {
        const int j = t.camrank[q];
        if (full) {
            t.pt[j] = a.tpt[q];
            t.ox[j] = a.ox[q];
            t.oy[j] = a.oy[q];
        }
        t.w[j] = make_float4(a.wx[q], a.wy[q], a.wxy[q], 0.f);
    }
    if (full && q < a.nch) t.ch_cam[q] = a.tcam[a.camq[a.ch_start[q]]];
}

template <typename R>
__global__ __launch_bounds__(64 * kMfChunkWaves) void k_mf_chunks(PcgArgs a)
{
    // a wavefront per chunk, kMfChunkWaves chunks per workgroup (fewer
    // workgroups to start, also when the PCG has stopped and they all return)
    if (a.S[a.j & 1].done) return;
    const int wv = threadIdx.x >> 6;
    const int ch = blockIdx.x * kMfChunkWaves + wv;
    if (ch >= a.nch) return;  // the whole wavefront; no workgroup barrier follows
    const int lane = threadIdx.x & 63;
    int b, n, k, first;
    if (a.ch4) {
        const int4 m = a.ch4[ch];
        b = m.x; n = m.y; k = m.z; first = m.w;
    } else {
        b = a.ch_start[ch];
        n = a.ch_start[ch + 1] - b;
        k = a.ch_cam[ch];
        first = ch == a.ch_off[k];
    }
    const R *Yr = (const R *)a.Yrec;
    // the entry's point, then the chunk's Y records (camera-major,
    // consecutive, read as whole lines) and the point's t behind them
    const int pq = a.cq_pt[b + min(lane, n - 1)];  // a chunk has at least one entry
    __shared__ u32x4 stg[kMfChunkWaves][RecSpan<R, kWRec>::LDS];
    RecSpan<R, kWRec> rw;
    span_issue<R, kWRec>(Yr, b, b + n - 1, rw, lane);
    const double *tq = a.tc + 3 * (size_t)pq;
    const double t0 = tq[0], t1 = tq[1], t2 = tq[2];
    R Yf[kWRec];  // Y as Bt^T P: the Bt pairs, then P = Bp V^-1 (2 x 3)
    span_transpose<R, kWRec>(rw, n, Yf, stg[wv], lane);
    double v[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) v[r] = 0.;
    if (lane < n) {
        const double p0 = (double)Yf[18] * t0 + (double)Yf[19] * t1 + (double)Yf[20] * t2;
        const double p1 = (double)Yf[21] * t0 + (double)Yf[22] * t1 + (double)Yf[23] * t2;
#pragma unroll
        for (int r = 0; r < 9; ++r) v[r] = (double)Yf[2 * r] * p0 + (double)Yf[2 * r + 1] * p1;
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) v[r] = wave_sum(v[r]);
    const double *dk = a.d + 9 * (size_t)k;
    if (lane < 9) {
        double vr = v[0];
#pragma unroll
        for (int r = 1; r < 9; ++r)
            if (lane == r) vr = v[r];
        a.cpart[9 * (size_t)ch + lane] = vr;
    }
    double share = 0.;
#pragma unroll
    for (int r = 0; r < 9; ++r) share -= dk[r] * v[r];
    if (first) {
        // (S_kk + D_k) d_k, row lane
        __shared__ double ed[kMfChunkWaves][9];
        if (lane < 9) {
            const double *S = a.Sd + 81 * (size_t)k + 9 * lane;
            const double *D = a.Dk + 81 * (size_t)k + 9 * lane;
            double y = 0.;
#pragma unroll
            for (int c = 0; c < 9; ++c) y += (S[c] + D[c]) * dk[c];
            a.Ad[9 * (size_t)k + lane] = y;
            ed[wv][lane] = y;
        }
        wave_lds_sync();
        double e = 0.;
#pragma unroll
        for (int r = 0; r < 9; ++r) e += dk[r] * ed[wv][r];
        share += e;
    }
    if (lane == 0) a.dpart[ch] = share;
}


// PCG step, part 1 (one workgroup per camera k): alpha = rho / d.S d from the
// tile partials; (S d)_k = yD_k + the run sums of row k + the column shares of
// column k, in a fixed order; x += alpha d, r -= alpha S d, h1 = C r; the
// camera's shares of |alpha S d|^2, r.(h1 - h) and r.h1
__device__ __forceinline__ void pcg_a_body(const PcgArgs &a, const PcgScalars &S, int k)
{
    __shared__ double sh[4];
    __shared__ double red[28][9];
    __shared__ double rs[9], p3[3][9];
    const int t = threadIdx.x;
    const int g = t / 9, r = t % 9;
    // the camera's own entries, loaded with the gathers below
    const size_t j = 9 * (size_t)k + (t < 9 ? t : 0);
    const double xj = a.x[j], dj = a.d[j], rj = a.r[j], hj = a.h[j];
    // the lane's row of the block-Jacobi inverse, in flight during the gathers
    double Crow[9];
    {
        const double *C = a.Cinv + 81 * (size_t)k + 9 * (t < 9 ? t : 0);
#pragma unroll
        for (int c = 0; c < 9; ++c) Crow[c] = C[c];
    }
    if (a.implicit) {
        // (S d)_k = (S_kk + D_k) d_k - the chunk sums of image k: group g
        // takes chunks g, g + 28, ... in order, the groups are added in order
        if (g < 28) {
            double acc = 0.;
            for (int ch = a.ch_off[k] + g; ch < a.ch_off[k + 1]; ch += 28) acc += a.cpart[9 * (size_t)ch + r];
            red[g][r] = (g == 0) ? a.Ad[9 * (size_t)k + r] - acc : -acc;
        }
    } else if (g < 28) {
    // row k: the row partials of its strip's chunks, in chunk order
    const int st = k / kSlab, gk = k - kSlab * st;
    const int c0 = a.chunk_off[st], nck = a.chunk_off[st + 1] - c0;
    {
        double acc = (g == 0) ? a.Ad[9 * (size_t)k + r] : 0.;
        for (int it = g; it < nck; it += 28) acc += a.rowpart[63 * (size_t)(c0 + it) + 9 * gk + r];
        // column k: the shares of its slabs, strip order; four loads in flight,
        // added in order
        const int q0 = a.cs_off[k], nq = a.cs_off[k + 1] - q0;
        const int *cl = a.cs_list + q0;
        int it = g;
        for (; it + 3 * 28 < nq; it += 4 * 28) {
            const int b0 = cl[it], b1 = cl[it + 28], b2 = cl[it + 56], b3 = cl[it + 84];
            const double v0 = a.colpart[9 * (size_t)b0 + r], v1 = a.colpart[9 * (size_t)b1 + r];
            const double v2 = a.colpart[9 * (size_t)b2 + r], v3 = a.colpart[9 * (size_t)b3 + r];
            acc += v0;
            acc += v1;
            acc += v2;
            acc += v3;
        }
        for (; it < nq; it += 28) acc += a.colpart[9 * (size_t)cl[it] + r];
        red[g][r] = acc;
    }
    }
    // d.S d after the gathers, so that its loads overlap theirs (its first
    // barrier also publishes red)
    const double dAd = sum_fixed(a.dsum, a.ndsum, sh);
    const double rho = S.rho;
    const double alpha = (fabs(rho) < 1.e-20 || fabs(rho) > 1.e20) ? 0. : rho / dAd;
    __syncthreads();
    if (t < 9) {
        double Ad = 0.;
        for (int gg = 0; gg < 28; ++gg) Ad += red[gg][t];
        a.x[j] = xj + alpha * dj;
        const double dr = alpha * Ad;
        const double rn = rj - dr;
        a.r[j] = rn;
        rs[t] = rn;
        p3[0][t] = dr * dr;
    }
    __syncthreads();
    if (t < 9) {
        double h1 = 0.;
#pragma unroll
        for (int c = 0; c < 9; ++c) h1 += Crow[c] * rs[c];
        a.h1[j] = h1;
        p3[1][t] = rs[t] * (h1 - hj);
        p3[2][t] = rs[t] * h1;
    }
    __syncthreads();
    if (t < 3) {
        double s = 0.;
        for (int q = 0; q < 9; ++q) s += p3[t][q];
        a.part[3 * (size_t)k + t] = s;
    }
    __syncthreads();  // LDS reused by the next camera of a persistent workgroup
}

__global__ __launch_bounds__(256) void k_pcg_a(PcgArgs a)
{
    const PcgScalars S = a.S[a.j & 1];
    if (S.done) return;
    pcg_a_body(a, S, blockIdx.x);
}

// u = N t for the corrected preconditioner (N: the off-diagonal blocks of S,
// both triangles), one workgroup per camera k: lanes in 28 groups of 9 take
// the row blocks (k, l) (S_kl t_l) and then the column blocks (l, k)
// (S_lk^T t_l) of camera k; the groups' sums are added in group order into u_k
__global__ __launch_bounds__(256) void k_offdiag_mv(PcgArgs a, const double *t_in, double *u)
{
    const PcgScalars S = a.S[a.j & 1];
    if (S.done) return;
    __shared__ double red[28][9];
    const int k = blockIdx.x, t = threadIdx.x;
    const int g = t / 9, r = t % 9;
    if (g < 28) {
        const int r0 = a.row_off[k], nr = a.row_off[k + 1] - r0;
        const int c0 = a.col_off[k], nl = a.col_off[k + 1] - c0;
        double acc = 0.;
        for (int e = g; e < nr + nl; e += 28) {
            if (e < nr) {
                const int b = r0 + e;
                const double *tl = t_in + 9 * (size_t)a.blk_l[b];
                double s = 0.;
#pragma unroll
                for (int c = 0; c < 9; ++c) s += a.Su[su_at(a.sm, b, r, c)] * tl[c];
                acc += s;
            } else {
                const int b = a.col_blk[c0 + e - nr];
                const double *tl = t_in + 9 * (size_t)a.blk_k[b];
                double s = 0.;
#pragma unroll
                for (int c = 0; c < 9; ++c) s += a.Su[su_at(a.sm, b, c, r)] * tl[c];
                acc += s;
            }
        }
        red[g][r] = acc;
    }
    __syncthreads();
    if (t < 9) {
        double s = 0.;
        for (int gg = 0; gg < 28; ++gg) s += red[gg][t];
        u[9 * (size_t)k + t] = s;
    }
}

// corrected preconditioner, h1 = t - w C_k u_k with t = C r already in tv (h
// at the set-up, h1 in an iteration) and u = N t from k_offdiag_mv (in Ad,
// free once k_pcg_a has used S d).  Set-up (init = 1): h = d = h1 and the
// share r.h; an iteration: h1 and the shares r.(h1 - h), r.h1 (the
// |alpha S d|^2 share of k_pcg_a stays).
__global__ __launch_bounds__(64) void k_pcg_c(PcgArgs a, int init)
{
    const PcgScalars S = a.S[a.j & 1];
    if (!init && S.done) return;
    __shared__ double us[9], p2[2][9];
    const int k = blockIdx.x, t = threadIdx.x;
    const float om = *a.omega;
    if (t < 9) us[t] = a.Ad[9 * (size_t)k + t];
    __syncthreads();
    if (t < 9) {
        const size_t j = 9 * (size_t)k + t;
        const double *C = a.Cinv + 81 * (size_t)k + 9 * t;
        double cu = 0.;
#pragma unroll
        for (int c = 0; c < 9; ++c) cu += C[c] * us[c];
        double *tv = init ? a.h : a.h1;
        const double h1 = tv[j] - (double)om * cu;
        const double rr = a.r[j];
        if (init) {
            a.h[j] = h1;
            a.d[j] = h1;
            p2[1][t] = rr * h1;
        } else {
            p2[0][t] = rr * (h1 - a.h[j]);
            p2[1][t] = rr * h1;
            a.h1[j] = h1;
        }
    }
    __syncthreads();
    if (t < 2 && !(init && t == 0)) {
        double sum = 0.;
        for (int q = 0; q < 9; ++q) sum += p2[t][q];
        a.part[3 * (size_t)k + 1 + t] = sum;
    }
}

// set-up of the corrected preconditioner, single workgroup: rho = r.h from
// the camera shares; when it is not positive (the correction is not positive
// definite along b) the weight is set to 0 and h = d = C b again
__global__ __launch_bounds__(1024) void k_pcg_rho(PcgArgs a)
{
    __shared__ double sh[16];
    double rh = 0.;
    for (int k = threadIdx.x; k < a.nc; k += blockDim.x) rh += a.part[3 * (size_t)k + 2];
    rh = block_sum(rh, sh);
    if (!(rh > 0.) || !isfinite(rh)) {
        __syncthreads();
        if (threadIdx.x == 0) *a.omega = 0.f;
        for (int k = threadIdx.x; k < a.nc; k += blockDim.x) apply_cinv(a.Cinv, a.r, a.h, k);
        __syncthreads();
        double s = 0.;
        for (int j = threadIdx.x; j < 9 * a.nc; j += blockDim.x) {
            a.d[j] = a.h[j];
            s += a.r[j] * a.h[j];
        }
        rh = block_sum(s, sh);
    }
    if (threadIdx.x == 0) {
        a.S[0].rho = rh;
        a.S[1].rho = rh;
    }
}

// PCG step, part 2: best iterate, stopping rule, beta, d = h1 + beta d, h = h1
// for the cameras of group grp (28 per group); group 0 writes the scalars of
// the next launch
__device__ __forceinline__ void pcg_b_body(const PcgArgs &a, const PcgScalars &S, PcgScalars *Sn, int grp)
{
    __shared__ double sh[3][4];
    const int G = a.nc;
    const int t = threadIdx.x;
    const int k = grp * 28 + t / 9;
    const bool on = (t < 252) && (k < a.nc);
    const size_t j = on ? 9 * (size_t)k + t % 9 : 0;
    // this thread's entries, loaded with the shares below
    const double xj = a.x[j], h1 = a.h1[j], dj = a.d[j];
    double m = 0., num = 0., rh = 0.;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        m += a.part[3 * g];
        num += a.part[3 * g + 1];
        rh += a.part[3 * g + 2];
    }
    // the three block sums in one pass (each in the order of block_sum)
    {
        m = wave_sum(m);
        num = wave_sum(num);
        rh = wave_sum(rh);
        const int w = t >> 6, nw = blockDim.x >> 6;
        if ((t & 63) == 0) {
            sh[0][w] = m;
            sh[1][w] = num;
            sh[2][w] = rh;
        }
        __syncthreads();
        m = num = rh = 0.;
        for (int i = 0; i < nw; ++i) {
            m += sh[0][i];
            num += sh[1][i];
            rh += sh[2][i];
        }
    }
    const int it = S.k;
    const double m0 = (it == 0) ? m : S.m0;
    const bool better = m < S.mold;
    if (better && on) a.best[j] = xj;
    PcgScalars z = S;
    z.m0 = m0;
    if (better) z.mold = m;
    if (m * 1.e5 < m0) {
        z.done = 1;
        z.iters = it + 1;
        if (grp == 0 && t == 0) *Sn = z;
        return;
    }
    double beta = (fabs(num) < 1.e-10) ? 0. : num / S.rho;
    if (beta < 0.) beta = 0.;
    if (on) {
        a.d[j] = h1 + beta * dj;
        a.h[j] = h1;
    }
    z.rho = rh;
    z.k = it + 1;
    if (it + 1 >= a.msize) {
        z.done = 1;
        z.iters = a.msize;
    }
    if (grp == 0 && t == 0) *Sn = z;
}

__global__ __launch_bounds__(256) void k_pcg_b(PcgArgs a)
{
    const PcgScalars S = a.S[a.j & 1];
    PcgScalars *Sn = a.S + ((a.j + 1) & 1);
    if (S.done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *Sn = S;
        return;
    }
    pcg_b_body(a, S, Sn, blockIdx.x);
}

// ---------------------------------------------------------------------------
// back-substitution, decode, cost
// ---------------------------------------------------------------------------

// dp_i = V_i^-1 (g_i - sum_q W_q^T dc_k) with the double camera solution
// (rbacore.cpp:1395-1410); Xn = X - dp (rbacore.cpp:547)
template <typename R, bool WF>
__global__ __launch_bounds__(256) void k_backsub(StepArgs a)
{
    // a workgroup per point group (first entry in a kPtGroup-entry window);
    // lane per entry u_q = W_q^T dc_cam(q) to LDS, then each point's lane
    // subtracts its entries from g_p in entry order
    __shared__ double su[256][3];
    const int t = threadIdx.x, c = blockIdx.x;
    const int4 grp = point_group(a.pg4, a.pgrp, a.toff, c);
    const int p0 = grp.x, p1 = grp.y, e0 = grp.z, e1 = grp.w;
    if (p0 >= p1) return;
    const int np = p1 - p0;
    const int i = p0 + t;
    int pb = 0, pe = 0;
    double t0 = 0., t1 = 0., t2 = 0.;
    if (t < np) {
        pb = a.toff[i];
        pe = a.toff[i + 1];
        t0 = a.gp[3 * (size_t)i];
        t1 = a.gp[3 * (size_t)i + 1];
        t2 = a.gp[3 * (size_t)i + 2];
    }
    // WF: W as its factors (kWRec elements, the matrix-free form), else the
    // kRec-element record
    constexpr int E = WF ? kWRec : kRec;
    const R *Wr = (const R *)a.Wrec;
    __shared__ u32x4 stg[4][RecSpan<R, E>::LDS];  // per wavefront
    const int lane = t & 63, wv = t >> 6;
    for (int base = e0; base < e1; base += 256) {
        const int q = base + t;
        const int cam = a.tcam[min(q, e1 - 1)];
        const int q0 = base + 64 * wv, nq = min(64, max(0, e1 - q0));
        RecSpan<R, E> rw;
        span_issue<R, E>(Wr, q0, e1 - 1, rw, lane);
        double xc[9];
        const double *xp = a.sol + 9 * (size_t)cam;
#pragma unroll
        for (int r = 0; r < 9; ++r) xc[r] = xp[r];
        R W[E];
        span_transpose<R, E>(rw, nq, W, stg[wv], lane);
        if (q < e1) {
            double u0 = 0., u1 = 0., u2 = 0.;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const double xr = xc[r];
                if constexpr (WF) {
                    u0 += (double)w_elem(W, r, 0) * xr;
                    u1 += (double)w_elem(W, r, 1) * xr;
                    u2 += (double)w_elem(W, r, 2) * xr;
                } else {
                    u0 += (double)W[r * 3] * xr;
                    u1 += (double)W[r * 3 + 1] * xr;
                    u2 += (double)W[r * 3 + 2] * xr;
                }
            }
            su[t][0] = u0;
            su[t][1] = u1;
            su[t][2] = u2;
        }
        __syncthreads();
        if (t < np) {
            const int lo = max(pb, base), hi = min(pe, base + 256);
            for (int x = lo; x < hi; ++x) {
                t0 -= su[x - base][0];
                t1 -= su[x - base][1];
                t2 -= su[x - base][2];
            }
        }
        __syncthreads();
    }
    if (t >= np) return;
    const double *Vi = a.Vinv + 9 * (size_t)i;
    bool bad = false;
    const double tt[3] = {t0, t1, t2};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double dp = Vi[r * 3] * tt[0] + Vi[r * 3 + 1] * tt[1] + Vi[r * 3 + 2] * tt[2];
        if (isnan(dp) || isinf(dp)) bad = true;
        a.Xn[3 * (size_t)i + r] = a.X[3 * (size_t)i + r] - dp;
    }
    if (bad) a.S->nan_flag = 1;
}

__device__ __forceinline__ void cayley(double a, double b, double c, double R[9])
{
    const double a2 = a * a, b2 = b * b, c2 = c * c;
    const double ab = a * b, ac = a * c, bc = b * c;
    double s = 4 + a2 + b2 + c2;
    if (fabs(s) < 1.e-20)
        s = (s < 0.) ? -1.e20 : 1.e20;
    else
        s = 1. / s;
    R[0] = s * (4 + a2 - b2 - c2);
    R[1] = s * (2 * ab - 4 * c);
    R[2] = s * (2 * ac + 4 * b);
    R[3] = s * (2 * ab + 4 * c);
    R[4] = s * (4 - a2 + b2 - c2);
    R[5] = s * (2 * bc - 4 * a);
    R[6] = s * (2 * ac - 4 * b);
    R[7] = s * (2 * bc + 4 * a);
    R[8] = s * (4 - a2 - b2 + c2);
}

// quaternion re-orthonormalisation (RectifyRotationMatrix3, rpba.h:876-913)
__device__ __forceinline__ void rectify(double R[9])
{
    double a0 = R[7] - R[5], a1 = R[2] - R[6], a2 = R[3] - R[1];
    if (fabs(a0) < 1.e-20) a0 = 1.e-20;
    if (fabs(a1) < 1.e-20) a1 = 1.e-20;
    if (fabs(a2) < 1.e-20) a2 = 1.e-20;
    const double al = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    a0 /= al;
    a1 /= al;
    a2 /= al;
    const double phi = atan2(al, R[0] + R[4] + R[8] - 1.);
    const double q0 = cos(phi * 0.5), sh = sin(phi * 0.5);
    const double q1 = a0 * sh, q2 = a1 * sh, q3 = a2 * sh;
    R[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    R[1] = 2 * (q1 * q2 - q0 * q3);
    R[2] = 2 * (q1 * q3 + q0 * q2);
    R[3] = 2 * (q2 * q1 + q0 * q3);
    R[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    R[5] = 2 * (q2 * q3 - q0 * q1);
    R[6] = 2 * (q3 * q1 - q0 * q2);
    R[7] = 2 * (q3 * q2 + q0 * q1);
    R[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
}

// theta_new = theta - (float) x_c; trial cameras (DecodeCameras,
// rbacore.cpp:1733-1800).  sol == nullptr decodes theta itself.
__global__ __launch_bounds__(64) void k_decode(StepArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nc) return;
    if (a.camflag && !(a.camflag[k] & 1)) return;
    const bool first = a.camflag ? (a.camflag[k] & 2) != 0 : k == 0;
    double th[9];
    bool bad = false;
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        double xs = 0.;
        if (a.sol) {
            const float xf = (float)a.sol[9 * (size_t)k + s];  // camera step cast to float (rbacore.cpp:1185)
            if (isnan(xf) || isinf(xf)) bad = true;
            xs = (double)xf;
        }
        th[s] = a.cam.th[9 * (size_t)k + s] - xs;
        a.cam.thn[9 * (size_t)k + s] = th[s];
    }
    if (bad) a.S->nan_flag = 1;
    const unsigned mask = a.cam.mask[k];
    if (!first) {
        double Ra[9], R[9];
        cayley(th[0], th[1], th[2], Ra);
        const double *Rb = a.cam.Rb + 9 * (size_t)k;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) R[r * 3 + c] = Ra[r * 3 + 0] * Rb[0 * 3 + c] + Ra[r * 3 + 1] * Rb[1 * 3 + c] + Ra[r * 3 + 2] * Rb[2 * 3 + c];
        rectify(R);
#pragma unroll
        for (int j = 0; j < 9; ++j) a.cam.Rn[9 * (size_t)k + j] = R[j];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double e = ((mask >> (3 + j)) & 1u) ? th[3 + j] : 0.;
            a.cam.Cn[3 * (size_t)k + j] = e + a.cam.Cb[3 * (size_t)k + j];
        }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j)
        if ((mask >> (6 + j)) & 1u) a.cam.In[3 * (size_t)k + j] = th[6 + j];
}

// per-entry residual at the trial state: float robust key (sig0rmv) and the
// FP64 cost term; per-workgroup partial sums
// the upper median's track-length factor of a key (Reweight2D rbacore.cpp:831-835)
__device__ __forceinline__ float redfac_of(int t)
{
    int pxs = t * 2;
    if (pxs < 4) pxs = 4;
    const float rf = (float)pxs / (float)(pxs - 3);
    return rf * rf;
}

__global__ __launch_bounds__(256) void k_cost(CostArgs a)
{
    __shared__ double sh[4];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    double cst = 0.;
    if (q < a.ntr) {
        const int k = a.tcam[q], p = a.tpt[q];
        CamV c;
        load_cam(a.cam, k, c);
        Ev ev;
        eval_obs(c, a.X[3 * p], a.X[3 * p + 1], a.X[3 * p + 2], a.ox[q], a.oy[q], a.defdiff, ev);
        float key = 1.e20f;
        if (ev.nan) {
            a.key[q] = key;
        } else {
            const float W0 = a.wx[q], W1 = a.wy[q], W2 = a.wxy[q];
            const float d1f = (float)ev.d1, d2f = (float)ev.d2;
            const float d11 = d1f * d1f, d12 = 2.f * d1f * d2f, d22 = d2f * d2f;
            if (a.robust)
                key = d11 * a.w0x[q] + d12 * a.w0xy[q] + d22 * a.w0y[q];
            else
                key = d11 * W0 + d12 * W2 + d22 * W1;
            if (a.write_key) a.key[q] = key;
            cst = ev.d1 * ev.d1 * (double)W0 + 2. * ev.d1 * ev.d2 * (double)W2 + ev.d2 * ev.d2 * (double)W1;
        }
        // the key as the upper median reads it, at the camera-major slot
        if (a.cm_key && (ev.nan || a.write_key)) a.cm_key[a.camrank[q]] = key * redfac_of(a.toff[p + 1] - a.toff[p]);
        if (a.terms) a.terms[q] = cst;
    }
    const double s = block_sum(cst, sh);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = s;
}

// sum of n partials in index order (single workgroup of 1024); out[0] = sum,
// optionally the integer partials too
__global__ __launch_bounds__(1024) void k_final_sum(const double *partial, int n, const int *ipartial, double *out,
                                                    int *iout)
{
    __shared__ double sh[16];
    double s = 0.;
    long long is = 0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        s += partial[j];
        if (ipartial) is += ipartial[j];
    }
    s = block_sum(s, sh);
    double si = block_sum((double)is, sh);
    if (threadIdx.x == 0) {
        out[0] = s;
        if (iout) iout[0] = (int)si;
    }
}

// ---------------------------------------------------------------------------
// robust reweighting (Reweight2D rbacore.cpp:823-887)
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned ord_bits(float f)
{
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// exact upper median (sorted[n/2]) of key*redfac per image by 4-pass radix
// select; s00[k] = 1/(median * 2.1981f), x0.001 while iflag > 1
constexpr int kMedianLds = 8192;  // entries of an image whose keys k_median keeps in LDS

__global__ __launch_bounds__(256) void k_median(RobustArgs a)
{
    const int k = blockIdx.x;
    const int b = a.camoff[k], e = a.camoff[k + 1], n = e - b;
    __shared__ unsigned hist[256];
    __shared__ unsigned s_prefix, s_rank;
    __shared__ unsigned ukey[kMedianLds];  // the image's ordered key bits (when it has at most kMedianLds entries)
    const bool staged = n <= kMedianLds;
    if (threadIdx.x == 0) {
        s_prefix = 0;
        s_rank = (unsigned)(n / 2);
    }
    unsigned maskbits = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        const unsigned prefix = s_prefix;
        for (int j = b + threadIdx.x; j < e; j += blockDim.x) {
            unsigned u;
            if (staged && shift != 24) {
                u = ukey[j - b];
            } else {
                float kv;
                if (a.cm_key) {
                    kv = a.cm_key[j];
                } else {
                    const int q = a.camq[j];
                    const int p = a.tpt[q];
                    kv = a.key[q] * redfac_of(a.toff[p + 1] - a.toff[p]);
                }
                u = ord_bits(kv);
                if (staged) ukey[j - b] = u;
            }
            if ((u & maskbits) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned cum = 0, rank = s_rank, sel = 255;
            for (unsigned d = 0; d < 256; ++d) {
                if (cum + hist[d] > rank) {
                    sel = d;
                    break;
                }
                cum += hist[d];
            }
            s_rank = rank - cum;
            s_prefix = prefix | (sel << shift);
        }
        maskbits |= 255u << shift;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned u = s_prefix;
        const unsigned raw = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
        const float med = __uint_as_float(raw);
        float s = 1.f / (med * 2.1981f);
        const int ifl = a.iflag_blk ? a.iflag_blk[a.camblk[k]] : a.iflag;
        if (ifl > 1) s *= 0.001f;
        a.s00[k] = s;
    }
}

__global__ __launch_bounds__(256) void k_reweight(RobustArgs a)
{
    __shared__ double sh[4];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    double ws = 0.;
    int nf = 0;
    if (q < a.ntr) {
        float ww = a.key[q] * a.s00[a.tcam[q]];
        const bool fl = ww > 9.f;  // kf, rpba.h:529
        a.flag[q] = fl ? 1 : 0;
        if (fl) {
            ww = 1.e-4f;
            nf = 1;
        } else
            ww = 1.f;
        float W0 = ww * a.w0x[q], W1 = ww * a.w0y[q], W2 = ww * a.w0xy[q];
        if (W0 + W1 < 2.e-8f) {
            W0 = W1 = 1.e-8f;
            W2 = 0.f;
        }
        a.wx[q] = W0;
        a.wy[q] = W1;
        a.wxy[q] = W2;
        if (a.cm_w) a.cm_w[a.camrank[q]] = make_float4(W0, W1, W2, 0.f);
        ws = (double)W0 + (double)W1;
    }
    const double s = block_sum(ws, sh);
    const double fsum = block_sum((double)nf, sh);
    if (threadIdx.x == 0) {
        a.partial[blockIdx.x] = s;
        a.ipartial[blockIdx.x] = (int)fsum;
    }
}

// sum of wx + wy over the entries (PrepareWeights2D's wsum, FP64)
__global__ __launch_bounds__(256) void k_wsum(int n, const float *wx, const float *wy, double *partial)
{
    __shared__ double sh[4];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const double v = (q < n) ? (double)wx[q] + (double)wy[q] : 0.;
    const double s = block_sum(v, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// StoreResultsIteration for the cameras (rbacore.cpp:895-906): P = [R | R C]
__global__ __launch_bounds__(64) void k_store_cams(CamPtrs cp, int nc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const double *R = cp.Rn + 9 * (size_t)k, *C = cp.Cn + 3 * (size_t)k;
    double *P = cp.Pst + 12 * (size_t)k;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) P[r * 4 + c] = R[r * 3 + c];
        P[r * 4 + 3] = R[r * 3 + 0] * C[0] + R[r * 3 + 1] * C[1] + R[r * 3 + 2] * C[2];
    }
    for (int j = 0; j < 3; ++j) cp.Ist[3 * (size_t)k + j] = cp.In[3 * (size_t)k + j];
    for (int j = 0; j < 9; ++j) cp.th[9 * (size_t)k + j] = cp.thn[9 * (size_t)k + j];
}

// ---------------------------------------------------------------------------
// Exteval (EvalLSA + CalcCorrPointMulti, rpbacore.cpp:882-941, 778-837)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_exteval(ExtArgs a)
{
    __shared__ double sh[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double contrib = 0.;
    int t = 0;
    if (i < a.npts) {
        const int b = a.toff[i], e = a.toff[i + 1];
        t = e - b;
        if (t > 1) {
            double ws = 0.;
            for (int q = b; q < e; ++q)
                ws += a.w ? (double)a.w[3 * (size_t)q] + (double)a.w[3 * (size_t)q + 1] : 2.;
            ws = ws / (double)(2 * t);
            const double X0 = a.pts[3 * i], X1 = a.pts[3 * i + 1], X2 = a.pts[3 * i + 2];
            double s = 0.;
            for (int q = b; q < e; ++q) {
                const int c = a.cam[q];
                const float *P = a.P + 12 * (size_t)c;
                double p1 = (double)P[0] * X0 + (double)P[1] * X1 + (double)P[2] * X2 + (double)P[3];
                double p2 = (double)P[4] * X0 + (double)P[5] * X1 + (double)P[6] * X2 + (double)P[7];
                double p3 = (double)P[8] * X0 + (double)P[9] * X1 + (double)P[10] * X2 + (double)P[11];
                if (fabs(p3) < 1.e-3)
                    p3 = (p3 > 0.) ? 1.e3 : -1.e3;
                else
                    p3 = 1. / p3;
                p1 *= p3;
                p2 *= p3;
                const double dist = p1 * p1 + p2 * p2;
                const double z = 1. + ((double)a.cd[2 * c] * dist + (double)a.cd[2 * c + 1] * dist * dist);
                p1 *= z;
                p2 *= z;
                p1 *= (double)a.cf[c];
                p2 *= (double)a.cf[c];
                const double dx = (double)a.xy[2 * (size_t)q] - p1, dy = (double)a.xy[2 * (size_t)q + 1] - p2;
                const float wx = a.w ? a.w[3 * (size_t)q] : 1.f, wy = a.w ? a.w[3 * (size_t)q + 1] : 1.f;
                const float wxy = a.w ? a.w[3 * (size_t)q + 2] : 0.f;
                s += dx * dx * wx + dy * dy * wy + 2 * dx * dy * wxy;
            }
            contrib = s / ws;
        } else
            t = 0;
    }
    const double sg = block_sum(contrib, sh);
    const double tt = block_sum((double)t, sh);
    const double cc = block_sum(t > 1 ? 1. : 0., sh);
    if (threadIdx.x == 0) {
        a.partial[blockIdx.x] = sg;
        a.tpartial[blockIdx.x] = (int)tt;
        a.cpartial[blockIdx.x] = (int)cc;
    }
}

// one workgroup: the partials in a fixed order, then sigma as EvalLSA forms it
// (rpbacore.cpp:923-938): dof = sum(2t - 3) - 6 (n - 2) - 5 - n * nraddp over
// points with t > 1, ihw2 = mean over images of ((w + h) / 2)^2 accumulated in
// float (rpbacore.cpp:87-90), sigma = sqrt(sum / dof * ihw2)
__global__ __launch_bounds__(1024) void k_exteval_finish(ExtArgs a, int nb, double *out)
{
    __shared__ double sh[16];
    double s = 0., ts = 0., cs = 0.;
    for (int j = threadIdx.x; j < nb; j += blockDim.x) {
        s += a.partial[j];
        ts += (double)a.tpartial[j];
        cs += (double)a.cpartial[j];
    }
    s = block_sum(s, sh);
    ts = block_sum(ts, sh);
    cs = block_sum(cs, sh);
    if (threadIdx.x == 0) {
        const long long tsum = (long long)ts, cnt = (long long)cs;
        const long long red = 2 * tsum - 3 * cnt, nobs = 2 * tsum;
        const long long dof = red - (long long)(a.ncams - 2) * 6 - 5 - (long long)a.ncams * a.nraddp;
        float ihw2 = 0.f;
        for (int k = 0; k < a.ncams; ++k) {
            const int w = a.img_w ? a.img_w[k] : 1, h = a.img_h ? a.img_h[k] : 1;
            ihw2 += (float)((h + w) * (h + w)) * 0.25f;
        }
        ihw2 /= (float)a.ncams;
        const double sg = s * (1. / (double)dof);
        out[0] = sqrt(sg * (double)ihw2);
        out[1] = (double)nobs;
        out[2] = (double)dof;
    }
}

// OutputLSAc + PointXXvOut on the device (rbacore.cpp:644-666, 1978-2010):
// the kept observations' flags in file order, and per workgroup the sum of
// squared residuals of the stored state (stored R|t, stored intrinsics,
// accepted points) for the RMS.  A fixed grid, so the sum is reproducible.
__global__ __launch_bounds__(256) void k_out_entries(int ntr, const int *tfo, const int *tcam, const int *tpt,
                                                     const float *ox, const float *oy, const double *Pst,
                                                     const double *Ist, const double *X, unsigned char *obs_kept,
                                                     double *partial)
{
    __shared__ double sh[4];
    double s = 0.;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ntr; q += gridDim.x * blockDim.x) {
        obs_kept[tfo[q]] = 1;
        const int k = tcam[q], i = tpt[q];
        const double *Pk = Pst + 12 * (size_t)k;
        const double *x = X + 3 * (size_t)i;
        const double e0 = Pk[0] * x[0] + Pk[1] * x[1] + Pk[2] * x[2] + Pk[3];
        const double e1 = Pk[4] * x[0] + Pk[5] * x[1] + Pk[6] * x[2] + Pk[7];
        const double e2 = Pk[8] * x[0] + Pk[9] * x[1] + Pk[10] * x[2] + Pk[11];
        const double u = e0 / e2, v = e1 / e2, rho = u * u + v * v;
        const double sc = 1. + Ist[3 * (size_t)k] * rho + Ist[3 * (size_t)k + 1] * rho * rho;
        const double r0 = (double)ox[q] - Ist[3 * (size_t)k + 2] * sc * u;
        const double r1 = (double)oy[q] - Ist[3 * (size_t)k + 2] * sc * v;
        s += r0 * r0 + r1 * r1;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_out_points(int npts, const int *nrpv, unsigned char *pt_kept)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npts; i += gridDim.x * blockDim.x) pt_kept[nrpv[i]] = 1;
}

// ---------------------------------------------------------------------------
// Whitened Jacobian records (precision 2).  With the observation weight
// W = L L^T (L lower: l11 = sqrt(wx), l21 = wxy / l11, l22 = sqrt(wy - l21^2))
// the rows of L^T Bc and L^T Bp are rounded once to R and stored per entry
// (kJRec values: 96 bytes in FP32).  Everything else is formed from them in
// FP64 (a product of two floats is exact in FP64):
//   V_i = sum Bp~^T Bp~, g_p = sum Bp~^T r~ (r~ = L^T r, FP64),
//   S_kk = sum_q Bc~^T K_q Bc~ with K_q = I - Bp~ V^-1 Bp~^T (2 x 2),
//   b_k = sum_q Bc~^T (r~ - Bp~ V^-1 g_p),
//   (S d)_k = sum_q Bc~^T (Bc~ d_k - Bp~ s_pt(q)) + delta_k d_k with
//   s_i = V_i^-1 sum_(q of i) Bp~^T Bc~ d_cam(q) and delta_k the damping of
//   S_kk's diagonal.
// The reduced camera system is then the Schur complement of the positive
// semi-definite sum J~^T J~ at every rounding of the records (plus the
// damping), so it stays positive definite and the PCG product is symmetric
// to FP64 rounding.  Records are point-major (Jpm, entry q at q) and
// camera-major (Jcm, camera-major entry j at j).
// ---------------------------------------------------------------------------

constexpr int kJRec = kJRecElems;  // Bc~ row 0 (9), row 1 (9), Bp~ row 0 (3), row 1 (3)

struct Whitener {
    double l11, l21, l22;
};
__device__ __forceinline__ Whitener whitener(float wx, float wy, float wxy)
{
    Whitener w;
    w.l11 = sqrt((double)wx);
    w.l21 = (double)wxy / w.l11;
    w.l22 = sqrt((double)wy - w.l21 * w.l21);
    return w;
}

template <typename R>
__device__ __forceinline__ void whiten_jac(const double Bc[2][9], const double Bp[2][3], const Whitener &w,
                                           R J[kJRec])
{
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        J[j] = (R)(w.l11 * Bc[0][j] + w.l21 * Bc[1][j]);
        J[9 + j] = (R)(w.l22 * Bc[1][j]);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        J[18 + j] = (R)(w.l11 * Bp[0][j] + w.l21 * Bp[1][j]);
        J[21 + j] = (R)(w.l22 * Bp[1][j]);
    }
}

template <typename R>
struct JSpan {
    static constexpr int NP = kJRec * (int)sizeof(R) / 16;  // pieces per record: 6 (FP32) or 12 (FP64)
    static constexpr int NPASS = sizeof(R) == 8 ? 2 : 1;   // LDS passes of the load transpose
    static constexpr int LDS = 64 * NP / NPASS;            // pieces of LDS per wavefront (load)
};

template <typename R>
__device__ __forceinline__ void jrec_pack(const R J[kJRec], u32x4 pc[JSpan<R>::NP])
{
#pragma unroll
    for (int p = 0; p < JSpan<R>::NP; ++p) {
        if constexpr (sizeof(R) == 8) {
            double2 d;
            d.x = J[2 * p];
            d.y = J[2 * p + 1];
            pc[p] = __builtin_bit_cast(u32x4, d);
        } else {
            float4 f;
            f.x = J[4 * p];
            f.y = J[4 * p + 1];
            f.z = J[4 * p + 2];
            f.w = J[4 * p + 3];
            pc[p] = __builtin_bit_cast(u32x4, f);
        }
    }
}

template <typename R>
__device__ __forceinline__ void jrec_unpack(const u32x4 pc[JSpan<R>::NP], R J[kJRec])
{
#pragma unroll
    for (int p = 0; p < JSpan<R>::NP; ++p) {
        if constexpr (sizeof(R) == 8) {
            const double2 d = __builtin_bit_cast(double2, pc[p]);
            J[2 * p] = d.x;
            J[2 * p + 1] = d.y;
        } else {
            const float4 f = __builtin_bit_cast(float4, pc[p]);
            J[4 * p] = f.x;
            J[4 * p + 1] = f.y;
            J[4 * p + 2] = f.z;
            J[4 * p + 3] = f.w;
        }
    }
}

// The wavefront's records q0 .. q0 + nq - 1 (consecutive from dst, NP pieces
// each) written as one contiguous span: each lane's record goes to the
// wavefront's LDS, then store instruction i writes piece 64 i + l of the
// span (whole lines).  Every lane of the wavefront calls it.
template <int NP>
__device__ __forceinline__ void jspan_store(u32x4 *dst, int nq, const u32x4 pc[NP], u32x4 *stg, int lane)
{
    if (lane < nq) {
#pragma unroll
        for (int p = 0; p < NP; ++p) stg[NP * lane + p] = pc[p];
    }
    wave_lds_sync();
    const int n = nq * NP;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int x = 64 * i + lane;
        if (x < n) dst[x] = stg[x];
    }
    wave_lds_sync();
}

// The records q0 .. q0 + 63 as one contiguous span of pieces (load
// instruction i, lane l: piece 64 i + l, clamped onto the last piece of
// record qlast, a valid record); jspan_transpose hands each record to its
// lane through the wavefront's LDS (NPASS passes of 64 / NPASS records).
// Every lane of the wavefront calls both; lanes past nq get nothing.
template <typename R>
__device__ __forceinline__ void jspan_issue(const void *recs, int q0, int qlast, u32x4 raw[JSpan<R>::NP], int lane)
{
    constexpr int NP = JSpan<R>::NP;
    const u32x4 *base = (const u32x4 *)recs + (size_t)q0 * NP;
    const int last = (qlast - q0 + 1) * NP - 1;
#pragma unroll
    for (int i = 0; i < NP; ++i) raw[i] = base[min(64 * i + lane, last)];
}

template <typename R>
__device__ __forceinline__ void jspan_transpose(const u32x4 raw[JSpan<R>::NP], int nq, R J[kJRec], u32x4 *stg, int lane)
{
    constexpr int NP = JSpan<R>::NP, NPASS = JSpan<R>::NPASS, IPP = NP / NPASS, RP = 64 / NPASS;
    u32x4 pc[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) pc[p] = u32x4{0u, 0u, 0u, 0u};  // lanes past nq get zeros (unused)
#pragma unroll
    for (int h = 0; h < NPASS; ++h) {
#pragma unroll
        for (int i = 0; i < IPP; ++i) stg[64 * i + lane] = raw[IPP * h + i];
        wave_lds_sync();
        const int rl = lane - RP * h;
        if (rl >= 0 && rl < RP && lane < nq) {
#pragma unroll
            for (int p = 0; p < NP; ++p) pc[p] = stg[NP * rl + p];
        }
        wave_lds_sync();
    }
    jrec_unpack<R>(pc, J);
}

// V, g_p and the point-major records: a workgroup per point group, lane per
// entry (Jacobians in FP64, whitened, rounded once to R, written as the
// wavefront's span; its V and g_p terms in FP64 from the rounded values to
// LDS), then each point's lane adds its entries in order; damping and V^-1
// as k_lin_points.
template <typename R>
__global__ __launch_bounds__(256) void k_lin_points_j(LinArgs a)
{
    constexpr int NP = JSpan<R>::NP;
    constexpr int SV = 256 * 12 / 2, ST = 4 * 64 * NP;  // pieces: per-entry terms; record staging
    __shared__ u32x4 lds[SV > ST ? SV : ST];
    double(*sv)[12] = reinterpret_cast<double(*)[12]>(lds);
    const int t = threadIdx.x, c = blockIdx.x, lane = t & 63, wv = t >> 6;
    const int4 grp = a.pg4[c];
    const int p0 = grp.x, p1 = grp.y, e0 = grp.z, e1 = grp.w;
    if (p0 >= p1) return;
    const int np = p1 - p0;
    const int i = p0 + t;
    int pb = 0, pe = 0;
    if (t < np) {
        pb = a.toff[i];
        pe = a.toff[i + 1];
    }
    double V[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0};
    u32x4 *Jpm = (u32x4 *)a.Jpm;
    for (int base = e0; base < e1; base += 256) {
        const int q = base + t;
        double term[12];
        u32x4 pc[NP];
        if (q < e1) {
            const int k = a.tcam[q], ip = a.tpt[q];
            CamV cv;
            load_cam(a.cam, k, cv);
            Ev ev;
            eval_obs(cv, a.X[3 * (size_t)ip], a.X[3 * (size_t)ip + 1], a.X[3 * (size_t)ip + 2], a.ox[q], a.oy[q], a.defdiff, ev);
            double Bc[2][9], Bp[2][3];
            jac_obs<double>(cv, ev, a.cam.mask[k], Bc, Bp);
            const Whitener w = whitener(a.wx[q], a.wy[q], a.wxy[q]);
            R J[kJRec];
            whiten_jac<R>(Bc, Bp, w, J);
            const double r0 = w.l11 * ev.d1 + w.l21 * ev.d2, r1 = w.l22 * ev.d2;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double b0 = (double)J[18 + r], b1 = (double)J[21 + r];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) term[r * 3 + cc] = b0 * (double)J[18 + cc] + b1 * (double)J[21 + cc];
                term[9 + r] = b0 * r0 + b1 * r1;
            }
            jrec_pack<R>(J, pc);
        }
        const int q0 = base + 64 * wv, nq = min(64, max(0, e1 - q0));
        if (nq > 0) jspan_store<NP>(Jpm + (size_t)q0 * NP, nq, pc, lds + 64 * NP * wv, lane);
        __syncthreads();
        if (q < e1) {
#pragma unroll
            for (int j = 0; j < 12; ++j) sv[t][j] = term[j];
        }
        __syncthreads();
        if (t < np) {
            const int lo = max(pb, base), hi = min(pe, base + 256);
            for (int x = lo; x < hi; ++x) {
#pragma unroll
                for (int j = 0; j < 9; ++j) V[j] += sv[x - base][j];
#pragma unroll
                for (int j = 0; j < 3; ++j) g[j] += sv[x - base][9 + j];
            }
        }
        __syncthreads();
    }
    if (t >= np) return;
    V[0] *= a.mult1;
    V[4] *= a.mult1;
    V[8] *= a.mult1;
    double Vi[9];
    inv3(V, Vi);  // V is symmetric to the bit, and so is its cofactor inverse
#pragma unroll
    for (int j = 0; j < 9; ++j) a.Vinv[9 * (size_t)i + j] = Vi[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) a.gp[3 * (size_t)i + j] = g[j];
    double2 *pr = (double2 *)(a.ptrec + 16 * (size_t)i);
    pr[0] = make_double2(a.X[3 * (size_t)i], a.X[3 * (size_t)i + 1]);
    pr[1] = make_double2(a.X[3 * (size_t)i + 2], Vi[0]);
    pr[2] = make_double2(Vi[1], Vi[2]);
    pr[3] = make_double2(Vi[4], Vi[5]);
    pr[4] = make_double2(Vi[8], g[0]);
    pr[5] = make_double2(g[1], g[2]);
    pr[6] = make_double2(0., 0.);
    pr[7] = make_double2(0., 0.);
}

// Camera blocks and the camera-major records: a workgroup per image, lane per
// camera-major entry (from the entry table; the point's X, V^-1 and g_p
// gathered): the same rounded whitened Jacobian as the point side, written as
// the wavefront's span, and the entry's terms Bc~^T K Bc~ (upper 45) and
// Bc~^T rho in FP64, added over the image (lanes, then the four wavefronts);
// S_kk = that sum with the diagonal damping x(1 + lambda), +1e-2
// (rbacore.cpp:1164-1170), the damping itself in dmp (1 on fixed slots),
// fixed slots an identity row / column with zero right-hand side.
template <typename R>
__global__ __launch_bounds__(256) void k_lin_cameras_j(LinArgs a)
{
    constexpr int NP = JSpan<R>::NP;
    const int k = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    __shared__ double red[4][54];
    __shared__ double Ssh[81];
    __shared__ double bsh[9];
    __shared__ u32x4 stg[4 * 64 * NP];
    const unsigned mask = a.cam.mask[k];
    CamV c;
    load_cam(a.cam, k, c);
    const int b = a.camoff[k], e = a.camoff[k + 1];
    double acc[54];
#pragma unroll
    for (int j = 0; j < 54; ++j) acc[j] = 0.;
    u32x4 *Jcm = (u32x4 *)a.Jcm;
    for (int j0 = b; j0 < e; j0 += 256) {
        const int j = j0 + t;
        u32x4 pc[NP];
        if (j < e) {
            // the point's line: X, V^-1 (symmetric), g_p
            const int p = a.cm_pt[j];
            const double2 *pr = (const double2 *)(a.ptrec + 16 * (size_t)p);
            const double2 r0v = pr[0], r1v = pr[1], r2v = pr[2], r3v = pr[3], r4v = pr[4], r5v = pr[5];
            const double Xp[3] = {r0v.x, r0v.y, r1v.x};
            const double Vi[9] = {r1v.y, r2v.x, r2v.y, r2v.x, r3v.x, r3v.y, r2v.y, r3v.y, r4v.x};
            const double gpv[3] = {r4v.y, r5v.x, r5v.y};
            Ev ev;
            const float4 wv = a.cm_w[j];
            eval_obs(c, Xp[0], Xp[1], Xp[2], a.cm_ox[j], a.cm_oy[j], a.defdiff, ev);
            double Bc[2][9], Bp[2][3];
            jac_obs<double>(c, ev, mask, Bc, Bp);
            const Whitener w = whitener(wv.x, wv.y, wv.z);
            R J[kJRec];
            whiten_jac<R>(Bc, Bp, w, J);
            const double r0 = w.l11 * ev.d1 + w.l21 * ev.d2, r1 = w.l22 * ev.d2;
            // P = Bp~ V^-1 (2 x 3), K = I - P Bp~^T, rho = r~ - P g_p
            double P[2][3];
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
                    P[s][cc] = (double)J[18 + 3 * s] * Vi[cc] + (double)J[19 + 3 * s] * Vi[3 + cc] +
                               (double)J[20 + 3 * s] * Vi[6 + cc];
            const double k00 = 1. - (P[0][0] * (double)J[18] + P[0][1] * (double)J[19] + P[0][2] * (double)J[20]);
            const double k01 = -(P[0][0] * (double)J[21] + P[0][1] * (double)J[22] + P[0][2] * (double)J[23]);
            const double k11 = 1. - (P[1][0] * (double)J[21] + P[1][1] * (double)J[22] + P[1][2] * (double)J[23]);
            const double rho0 = r0 - (P[0][0] * gpv[0] + P[0][1] * gpv[1] + P[0][2] * gpv[2]);
            const double rho1 = r1 - (P[1][0] * gpv[0] + P[1][1] * gpv[1] + P[1][2] * gpv[2]);
            int n = 0;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const double a0 = (double)J[r], a1 = (double)J[9 + r];
                const double t0 = a0 * k00 + a1 * k01, t1 = a0 * k01 + a1 * k11;
#pragma unroll
                for (int cc = r; cc < 9; ++cc) acc[n++] += t0 * (double)J[cc] + t1 * (double)J[9 + cc];
                acc[45 + r] += a0 * rho0 + a1 * rho1;
            }
            jrec_pack<R>(J, pc);
        }
        const int q0 = j0 + 64 * wv, nq = min(64, max(0, e - q0));
        if (nq > 0) jspan_store<NP>(Jcm + (size_t)q0 * NP, nq, pc, stg + 64 * NP * wv, lane);
    }
#pragma unroll
    for (int j = 0; j < 54; ++j) {
        const double s = wave_sum(acc[j]);
        if (lane == 0) red[wv][j] = s;
    }
    __syncthreads();
    if (t < 81) {
        const int r = t / 9, cc = t % 9;
        const int rr = r < cc ? r : cc, c2 = r < cc ? cc : r;
        const int n = rr * 9 - rr * (rr - 1) / 2 + (c2 - rr);  // index in the packed upper triangle
        const double s0 = red[0][n] + red[1][n] + red[2][n] + red[3][n];
        double s = s0;
        if (r == cc) {
            s *= a.mult1;
            s += 1.e-2;
        }
        const bool fr = (mask >> r) & 1u, fc = (mask >> cc) & 1u;
        if (!fr || !fc) s = (r == cc) ? 1. : 0.;
        Ssh[t] = s;
        if (r == cc) a.dmp[9 * (size_t)k + r] = fr ? s - s0 : 1.;
    } else if (t >= 128 && t < 137) {
        const int r = t - 128;
        double s = red[0][45 + r] + red[1][45 + r] + red[2][45 + r] + red[3][45 + r];
        if (!((mask >> r) & 1u)) s = 0.;
        bsh[r] = s;
    }
    __syncthreads();
    if (t < 81) a.Sd[81 * (size_t)k + t] = Ssh[t];
    if (t < 9) a.bvec[9 * (size_t)k + t] = bsh[t];
}

// Matrix-free S d through the whitened records, point side: a workgroup per
// point group, lane per entry u_q = Bp~^T (Bc~ d_cam(q)) (the records as the
// wavefront's span, the image index and d behind them) to LDS; each point's
// lane adds its entries in order and stores s_i = V_i^-1 t_i (tc).
template <typename R>
__global__ __launch_bounds__(256) void k_mf_points_j(PcgArgs a)
{
    if (a.S[a.j & 1].done) return;
    __shared__ double su[256][3];
    __shared__ u32x4 stg[4][JSpan<R>::LDS];
    const int t = threadIdx.x, c = blockIdx.x, lane = t & 63, wv = t >> 6;
    const int4 g = a.pg4[c];
    const int p0 = g.x, p1 = g.y, e0 = g.z, e1 = g.w;
    if (p0 >= p1) return;
    const int np = p1 - p0;
    // the point's entry range (and with FP32 records its V^-1, which FP64
    // records have no registers to spare for), in flight with the records
    int pb = 0, pe = 0;
    double Vi[9];
    const int pt = min(p0 + t, p1 - 1);
    pb = a.toff[pt];
    pe = a.toff[pt + 1];
    if constexpr (sizeof(R) == 4) {
        const double *vp = a.Vinv + 9 * (size_t)pt;
#pragma unroll
        for (int r = 0; r < 9; ++r) Vi[r] = vp[r];
    }
    double t0 = 0., t1 = 0., t2 = 0.;
    for (int base = e0; base < e1; base += 256) {
        const int q = base + t;
        const int cam = a.tcam[min(q, e1 - 1)];
        const int q0 = base + 64 * wv, nq = min(64, max(0, e1 - q0));
        u32x4 raw[JSpan<R>::NP];
        jspan_issue<R>(a.Jpm, q0, e1 - 1, raw, lane);
        double dv[9];
        const double *dp = a.d + 9 * (size_t)cam;
#pragma unroll
        for (int r = 0; r < 9; ++r) dv[r] = dp[r];
        R J[kJRec];
        jspan_transpose<R>(raw, nq, J, stg[wv], lane);
        if (q < e1) {
            double v0 = 0., v1 = 0.;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                v0 += (double)J[r] * dv[r];
                v1 += (double)J[9 + r] * dv[r];
            }
            su[t][0] = (double)J[18] * v0 + (double)J[21] * v1;
            su[t][1] = (double)J[19] * v0 + (double)J[22] * v1;
            su[t][2] = (double)J[20] * v0 + (double)J[23] * v1;
        }
        __syncthreads();
        if (t < np) {
            const int lo = max(pb, base), hi = min(pe, base + 256);
            for (int x = lo; x < hi; ++x) {
                t0 += su[x - base][0];
                t1 += su[x - base][1];
                t2 += su[x - base][2];
            }
        }
        __syncthreads();
    }
    if (t < np) {
        if constexpr (sizeof(R) == 8) {
            const double *vp = a.Vinv + 9 * (size_t)pt;
#pragma unroll
            for (int r = 0; r < 9; ++r) Vi[r] = vp[r];
        }
        double *o = a.tc + 3 * (size_t)(p0 + t);
#pragma unroll
        for (int r = 0; r < 3; ++r) o[r] = Vi[3 * r] * t0 + Vi[3 * r + 1] * t1 + Vi[3 * r + 2] * t2;
    }
}

// Camera side: a wavefront per camera chunk, lane per entry
// Bc~^T (Bc~ d_k - Bp~ s_pt(q)) (camera-major records as the span, the
// entry's point and its s gathered behind them), summed over the chunk in
// lane order and stored negated in cpart (k_pcg_a subtracts the chunk sums
// from Ad); the image's first chunk stores Ad_k = dmp_k d_k.  Shares of
// d.S d as k_mf_chunks.
template <typename R>
__global__ __launch_bounds__(64 * kMfChunkWaves) void k_mf_chunks_j(PcgArgs a)
{
    if (a.S[a.j & 1].done) return;
    const int wv = threadIdx.x >> 6;
    const int ch = blockIdx.x * kMfChunkWaves + wv;
    if (ch >= a.nch) return;  // the whole wavefront; no workgroup barrier follows
    const int lane = threadIdx.x & 63;
    const int4 m = a.ch4[ch];
    const int b = m.x, n = m.y, k = m.z;
    const int pq = a.cq_pt[b + min(lane, n - 1)];  // a chunk has at least one entry
    __shared__ u32x4 stg[kMfChunkWaves][JSpan<R>::LDS];
    u32x4 raw[JSpan<R>::NP];
    jspan_issue<R>(a.Jcm, b, b + n - 1, raw, lane);
    const double *sq = a.tc + 3 * (size_t)pq;
    const double s0 = sq[0], s1 = sq[1], s2 = sq[2];
    const double *dk = a.d + 9 * (size_t)k;
    double dv[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) dv[r] = dk[r];
    R J[kJRec];
    jspan_transpose<R>(raw, n, J, stg[wv], lane);
    double v[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) v[r] = 0.;
    if (lane < n) {
        double v0 = 0., v1 = 0.;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            v0 += (double)J[r] * dv[r];
            v1 += (double)J[9 + r] * dv[r];
        }
        const double w0 = v0 - ((double)J[18] * s0 + (double)J[19] * s1 + (double)J[20] * s2);
        const double w1 = v1 - ((double)J[21] * s0 + (double)J[22] * s1 + (double)J[23] * s2);
#pragma unroll
        for (int r = 0; r < 9; ++r) v[r] = -((double)J[r] * w0 + (double)J[9 + r] * w1);
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) v[r] = wave_sum(v[r]);
    if (lane < 9) {
        double vr = v[0];
#pragma unroll
        for (int r = 1; r < 9; ++r)
            if (lane == r) vr = v[r];
        a.cpart[9 * (size_t)ch + lane] = vr;
    }
    double share = 0.;
#pragma unroll
    for (int r = 0; r < 9; ++r) share -= dv[r] * v[r];
    if (m.w) {
        double e = 0.;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const double y = a.dmp[9 * (size_t)k + r] * dv[r];
            e += dv[r] * y;
            if (lane == r) a.Ad[9 * (size_t)k + r] = y;
        }
        share += e;
    }
    if (lane == 0) a.dpart[ch] = share;
}

// back-substitution through the whitened records: dp_i = V_i^-1 (g_p -
// sum_q Bp~^T (Bc~ dc_k)); Xn = X - dp (rbacore.cpp:1395-1410, 547)
template <typename R>
__global__ __launch_bounds__(256) void k_backsub_j(StepArgs a)
{
    __shared__ double su[256][3];
    __shared__ u32x4 stg[4][JSpan<R>::LDS];
    const int t = threadIdx.x, c = blockIdx.x, lane = t & 63, wv = t >> 6;
    const int4 g = a.pg4[c];
    const int p0 = g.x, p1 = g.y, e0 = g.z, e1 = g.w;
    if (p0 >= p1) return;
    const int np = p1 - p0;
    const int i = p0 + t;
    int pb = 0, pe = 0;
    double t0 = 0., t1 = 0., t2 = 0.;
    if (t < np) {
        pb = a.toff[i];
        pe = a.toff[i + 1];
        t0 = a.gp[3 * (size_t)i];
        t1 = a.gp[3 * (size_t)i + 1];
        t2 = a.gp[3 * (size_t)i + 2];
    }
    for (int base = e0; base < e1; base += 256) {
        const int q = base + t;
        const int cam = a.tcam[min(q, e1 - 1)];
        const int q0 = base + 64 * wv, nq = min(64, max(0, e1 - q0));
        u32x4 raw[JSpan<R>::NP];
        jspan_issue<R>(a.Jpm, q0, e1 - 1, raw, lane);
        double xc[9];
        const double *xp = a.sol + 9 * (size_t)cam;
#pragma unroll
        for (int r = 0; r < 9; ++r) xc[r] = xp[r];
        R J[kJRec];
        jspan_transpose<R>(raw, nq, J, stg[wv], lane);
        if (q < e1) {
            double v0 = 0., v1 = 0.;
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                v0 += (double)J[r] * xc[r];
                v1 += (double)J[9 + r] * xc[r];
            }
            su[t][0] = (double)J[18] * v0 + (double)J[21] * v1;
            su[t][1] = (double)J[19] * v0 + (double)J[22] * v1;
            su[t][2] = (double)J[20] * v0 + (double)J[23] * v1;
        }
        __syncthreads();
        if (t < np) {
            const int lo = max(pb, base), hi = min(pe, base + 256);
            for (int x = lo; x < hi; ++x) {
                t0 -= su[x - base][0];
                t1 -= su[x - base][1];
                t2 -= su[x - base][2];
            }
        }
        __syncthreads();
    }
    if (t >= np) return;
    const double *Vi = a.Vinv + 9 * (size_t)i;
    bool bad = false;
    const double tt[3] = {t0, t1, t2};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double dp = Vi[r * 3] * tt[0] + Vi[r * 3 + 1] * tt[1] + Vi[r * 3 + 2] * tt[2];
        if (isnan(dp) || isinf(dp)) bad = true;
        a.Xn[3 * (size_t)i + r] = a.X[3 * (size_t)i + r] - dp;
    }
    if (bad) a.S->nan_flag = 1;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_lin_points(const LinArgs &a, hipStream_t st)
{
    if (a.npts <= 0 || a.npgrp <= 0) return;
    // the matrix-free form: W (its factors) written by the point kernel;
    // else W and Y by k_lin_records, which needs V^-1
    const unsigned g = nblk(a.ntr, 128);
    if (a.rec32) {
        if (a.ycm) {
            k_lin_points<float, true><<<a.npgrp, 256, 0, st>>>(a);
        } else {
            k_lin_points<float, false><<<a.npgrp, 256, 0, st>>>(a);
            k_lin_records<float, true><<<g, 128, 0, st>>>(a);
        }
    } else {
        if (a.ycm) {
            k_lin_points<double, true><<<a.npgrp, 256, 0, st>>>(a);
        } else {
            k_lin_points<double, false><<<a.npgrp, 256, 0, st>>>(a);
            k_lin_records<double, true><<<g, 128, 0, st>>>(a);
        }
    }
}
void launch_schur_pairs(const SchurArgs &s, hipStream_t st)
{
    if (s.nblk <= 0 || s.nslab <= 0) return;
    // four wavefronts (slabs) per workgroup, at most 8 x 16384 workgroups
    const int per = (s.nslab + 7) / 8;
    const unsigned g = 8 * (unsigned)std::min((per + 3) / 4, 16384);
    if (s.rec32)
        k_schur_slabs<float><<<g, 256, 0, st>>>(s);
    else
        k_schur_slabs<double><<<g, 256, 0, st>>>(s);
}
void launch_schur_diag(const LinArgs &a, hipStream_t st)
{
    if (a.nch <= 0) return;
    if (a.ycm) {
        if (a.rec32)
            k_lin_chunks<float><<<a.nch, 64, 0, st>>>(a);
        else
            k_lin_chunks<double><<<a.nch, 64, 0, st>>>(a);
        return;
    }
    if (a.rec32)
        k_schur_diag<float><<<a.nch, 64, 0, st>>>(a);
    else
        k_schur_diag<double><<<a.nch, 64, 0, st>>>(a);
}
void launch_lin_cameras(const LinArgs &a, hipStream_t st)
{
    if (a.rec32)
        k_lin_cameras<float><<<a.nc, 256, 0, st>>>(a);
    else
        k_lin_cameras<double><<<a.nc, 256, 0, st>>>(a);
}
void launch_lin_points_j(const LinArgs &a, hipStream_t st)
{
    if (a.npts <= 0 || a.npgrp <= 0) return;
    if (a.rec32)
        k_lin_points_j<float><<<a.npgrp, 256, 0, st>>>(a);
    else
        k_lin_points_j<double><<<a.npgrp, 256, 0, st>>>(a);
}
void launch_lin_cameras_j(const LinArgs &a, hipStream_t st)
{
    if (a.rec32)
        k_lin_cameras_j<float><<<a.nc, 256, 0, st>>>(a);
    else
        k_lin_cameras_j<double><<<a.nc, 256, 0, st>>>(a);
}
void launch_bj_inverse(int nc, const double *Sd, double *Cinv, hipStream_t st)
{
    k_bj_inverse<<<nblk(nc, 64), 64, 0, st>>>(nc, Sd, Cinv);
}
void launch_pcg_init(const PcgArgs &a, hipStream_t st)
{
    const int nb = (int)nblk(a.nc, 256);  // part holds 3 nc doubles: room for the shares
    k_pcg_init<<<nb, 256, 0, st>>>(a);
    k_pcg_init_fin<<<1, 1024, 0, st>>>(a, nb);
}
// compute units of the current device (persistent SpMV grid)
static int device_cus()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    return n;
}
// SpMV workgroups: four per CU (persistent over the tiles)
// workgroups of the SpMV: as many as the chip holds at once (its occupancy
// per CU times the CUs), unless RPBA_SPMV_WPC (1..16) sets the count per CU
// for a launch-shape sweep.  The grid changes the order of the d.S d partial
// sums, not the products.
int spmv_grid(int nblk)
{
    (void)nblk;
    static int per_cu = 0;
    if (!per_cu) {
        const char *e = getenv("RPBA_SPMV_WPC");
        const int v = e ? atoi(e) : 0;
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_spmv_slabs, 256, 0) != hipSuccess || occ < 1) occ = 4;
        per_cu = (v >= 1 && v <= 16) ? v : occ;
    }
    return per_cu * device_cus();
}
void launch_spmv(const PcgArgs &a, hipStream_t st)
{
    if (!a.implicit) {
        k_spmv_slabs<<<a.ndpart, 256, 0, st>>>(a, a.chunk_strip, a.strip_s0, a.chunk_off, a.slab_bs, a.slab_mask,
                                               a.slab_l, a.d);
        return;
    }
    if (a.npgrp <= 0 || a.nch <= 0) return;
    if (a.jform) {
        if (a.rec32) {
            k_mf_points_j<float><<<a.npgrp, 256, 0, st>>>(a);
            k_mf_chunks_j<float><<<(a.nch + kMfChunkWaves - 1) / kMfChunkWaves, 64 * kMfChunkWaves, 0, st>>>(a);
        } else {
            k_mf_points_j<double><<<a.npgrp, 256, 0, st>>>(a);
            k_mf_chunks_j<double><<<(a.nch + kMfChunkWaves - 1) / kMfChunkWaves, 64 * kMfChunkWaves, 0, st>>>(a);
        }
    } else if (a.rec32) {
        k_mf_points<float><<<a.npgrp, 256, 0, st>>>(a);
        k_mf_chunks<float><<<(a.nch + kMfChunkWaves - 1) / kMfChunkWaves, 64 * kMfChunkWaves, 0, st>>>(a);
    } else {
        k_mf_points<double><<<a.npgrp, 256, 0, st>>>(a);
        k_mf_chunks<double><<<(a.nch + kMfChunkWaves - 1) / kMfChunkWaves, 64 * kMfChunkWaves, 0, st>>>(a);
    }
    k_mf_dsum<<<kMfDsumGroups, 256, 0, st>>>(a);
}
int pcg_groups(int nc) { return (nc + 27) / 28; }
void launch_cm_table(const LinArgs &a, const CmTable &t, bool full, hipStream_t st)
{
    const int n = std::max(a.ntr, a.nch);
    if (n > 0) k_cm_table<<<nblk(n, 256), 256, 0, st>>>(a, t, full ? 1 : 0);
}
void launch_pcg_init_w(const PcgArgs &a, hipStream_t st)
{
    PcgArgs b = a;
    b.j = 0;
    k_offdiag_mv<<<a.nc, 256, 0, st>>>(b, a.h, a.Ad);  // N (C b)
    k_pcg_c<<<a.nc, 64, 0, st>>>(b, 1);
    k_pcg_rho<<<1, 1024, 0, st>>>(b);
}
void launch_pcg_update_w(const PcgArgs &a, hipStream_t st)
{
    k_pcg_a<<<a.nc, 256, 0, st>>>(a);
    k_offdiag_mv<<<a.nc, 256, 0, st>>>(a, a.h1, a.Ad);  // N (C r)
    k_pcg_c<<<a.nc, 64, 0, st>>>(a, 0);
    k_pcg_b<<<pcg_groups(a.nc), 256, 0, st>>>(a);
}
void launch_pcg_update(const PcgArgs &a, hipStream_t st)
{
    k_pcg_a<<<a.nc, 256, 0, st>>>(a);
    k_pcg_b<<<pcg_groups(a.nc), 256, 0, st>>>(a);
}
void launch_backsub(const StepArgs &a, hipStream_t st)
{
    if (a.npts <= 0 || a.npgrp <= 0) return;
    if (a.jform) {
        if (a.rec32)
            k_backsub_j<float><<<a.npgrp, 256, 0, st>>>(a);
        else
            k_backsub_j<double><<<a.npgrp, 256, 0, st>>>(a);
    } else if (a.wfac) {
        if (a.rec32)
            k_backsub<float, true><<<a.npgrp, 256, 0, st>>>(a);
        else
            k_backsub<double, true><<<a.npgrp, 256, 0, st>>>(a);
    } else if (a.rec32) {
        k_backsub<float, false><<<a.npgrp, 256, 0, st>>>(a);
    } else {
        k_backsub<double, false><<<a.npgrp, 256, 0, st>>>(a);
    }
}
void launch_decode(const StepArgs &a, hipStream_t st) { k_decode<<<nblk(a.nc, 64), 64, 0, st>>>(a); }
int cost_blocks(int ntr) { return (int)nblk(ntr > 0 ? ntr : 1, 256); }
void launch_cost(const CostArgs &a, double *out, hipStream_t st)
{
    const int nb = cost_blocks(a.ntr);
    k_cost<<<nb, 256, 0, st>>>(a);
    k_final_sum<<<1, 1024, 0, st>>>(a.partial, nb, nullptr, out, nullptr);
}
void launch_reweight(const RobustArgs &a, double *wsum_out, int *nflag_out, hipStream_t st)
{
    k_median<<<a.nc, 256, 0, st>>>(a);
    const int nb = cost_blocks(a.ntr);
    k_reweight<<<nb, 256, 0, st>>>(a);
    k_final_sum<<<1, 1024, 0, st>>>(a.partial, nb, a.ipartial, wsum_out, nflag_out);
}
void launch_wsum(int n, const float *wx, const float *wy, double *partial, double *out, hipStream_t st)
{
    const int nb = cost_blocks(n);
    k_wsum<<<nb, 256, 0, st>>>(n, wx, wy, partial);
    k_final_sum<<<1, 1024, 0, st>>>(partial, nb, nullptr, out, nullptr);
}
void launch_store_cams(const CamPtrs &cp, int nc, hipStream_t st) { k_store_cams<<<nblk(nc, 64), 64, 0, st>>>(cp, nc); }
int out_groups(int ntr) { return (int)std::min<long long>(nblk(ntr > 0 ? ntr : 1, 256), 2048); }
void launch_output(int ntr, int npts, const int *tfo, const int *tcam, const int *tpt, const int *nrpv, const float *ox,
                   const float *oy, const double *Pst, const double *Ist, const double *X, unsigned char *obs_kept,
                   unsigned char *pt_kept, double *partial, double *ss_out, hipStream_t st)
{
    const int g = out_groups(ntr);
    k_out_entries<<<g, 256, 0, st>>>(ntr, tfo, tcam, tpt, ox, oy, Pst, Ist, X, obs_kept, partial);
    k_final_sum<<<1, 1024, 0, st>>>(partial, g, nullptr, ss_out, nullptr);
    if (npts > 0) k_out_points<<<(unsigned)std::min<long long>(nblk(npts, 256), 2048), 256, 0, st>>>(npts, nrpv, pt_kept);
}
int exteval_groups(int npts) { return (int)nblk(npts > 0 ? npts : 1, 256); }
void launch_exteval(const ExtArgs &a, double *out, hipStream_t st)
{
    const int nb = exteval_groups(a.npts);
    k_exteval<<<nb, 256, 0, st>>>(a);
    k_exteval_finish<<<1, 1024, 0, st>>>(a, nb, out);
}

// ---------------------------------------------------------------------------
// batched independent blocks
// ---------------------------------------------------------------------------

// per (block, slice): sums over one of kSlices equal parts of the block's
// entry range, then per block the slices in order
constexpr int kSlices = 32;

__global__ __launch_bounds__(256) void k_block_sums(BatchArgs b, const double *terms, const float *wa, const float *wb,
                                                    const unsigned char *flag, double *part)
{
    __shared__ double sh[4];
    const int blk = blockIdx.y, sl = blockIdx.x;
    const long long e0 = b.eoff[blk], n = b.eoff[blk + 1] - e0;
    const int q0 = (int)(e0 + n * sl / kSlices), q1 = (int)(e0 + n * (sl + 1) / kSlices);
    double s0 = 0., s1 = 0., s2 = 0.;
    for (int q = q0 + (int)threadIdx.x; q < q1; q += 256) {
        if (terms) s0 += terms[q];
        if (wa) s1 += (double)wa[q] + (double)wb[q];
        if (flag) s2 += (double)flag[q];
    }
    s0 = block_sum(s0, sh);
    s1 = block_sum(s1, sh);
    s2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        double *o = part + 3 * ((size_t)blk * kSlices + sl);
        o[0] = s0;
        o[1] = s1;
        o[2] = s2;
    }
}

__global__ __launch_bounds__(64) void k_block_sums_fin(int nb, const double *part, double *out)
{
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nb) return;
    double s0 = 0., s1 = 0., s2 = 0.;
    for (int sl = 0; sl < kSlices; ++sl) {
        const double *o = part + 3 * ((size_t)blk * kSlices + sl);
        s0 += o[0];
        s1 += o[1];
        s2 += o[2];
    }
    out[3 * blk] = s0;
    out[3 * blk + 1] = s1;
    out[3 * blk + 2] = s2;
}

__global__ __launch_bounds__(256) void k_block_nan(BatchArgs b, const double *sol, const double *Xn, int *part)
{
    __shared__ int any;
    const int blk = blockIdx.y, sl = blockIdx.x;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    int bad = 0;
    const long long c0 = 9LL * b.coff[blk], nc = 9LL * b.coff[blk + 1] - c0;
    for (long long j = c0 + nc * sl / kSlices + threadIdx.x; j < c0 + nc * (sl + 1) / kSlices; j += 256) {
        const float xf = (float)sol[j];
        if (isnan(xf) || isinf(xf)) bad = 1;
    }
    const long long p0 = 3LL * b.poff[blk], np = 3LL * b.poff[blk + 1] - p0;
    for (long long j = p0 + np * sl / kSlices + threadIdx.x; j < p0 + np * (sl + 1) / kSlices; j += 256)
        if (!isfinite(Xn[j])) bad = 1;
    if (bad) any = 1;  // every writer stores the same value
    __syncthreads();
    if (threadIdx.x == 0) part[(size_t)blk * kSlices + sl] = any;
}

__global__ __launch_bounds__(64) void k_block_nan_fin(int nb, const int *part, int *nanb)
{
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nb) return;
    int any = 0;
    for (int sl = 0; sl < kSlices; ++sl) any |= part[(size_t)blk * kSlices + sl];
    nanb[blk] = any;
}

__global__ __launch_bounds__(256) void k_batch_store(BatchArgs b, const unsigned char *accept, CamPtrs cp,
                                                     const double *Xn, double *X, const unsigned char *flag,
                                                     unsigned char *flagst)
{
    // grid-stride over cameras, points and entries: the launch grid is capped
    const int j0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (int j = j0; j < b.nc; j += stride) {
        if (!accept[b.camblk[j]]) continue;
        const double *R = cp.Rn + 9 * (size_t)j, *C = cp.Cn + 3 * (size_t)j;
        double *P = cp.Pst + 12 * (size_t)j;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) P[r * 4 + c] = R[r * 3 + c];
            P[r * 4 + 3] = R[r * 3 + 0] * C[0] + R[r * 3 + 1] * C[1] + R[r * 3 + 2] * C[2];
        }
        for (int s = 0; s < 3; ++s) cp.Ist[3 * (size_t)j + s] = cp.In[3 * (size_t)j + s];
        for (int s = 0; s < 9; ++s) cp.th[9 * (size_t)j + s] = cp.thn[9 * (size_t)j + s];
    }
    for (int j = j0; j < b.npts; j += stride)
        if (accept[b.ptblk[j]])
            for (int s = 0; s < 3; ++s) X[3 * (size_t)j + s] = Xn[3 * (size_t)j + s];
    for (int q = j0; q < b.ntr; q += stride)
        if (accept[b.ptblk[b.tpt[q]]]) flagst[q] = flag[q];
}

__global__ __launch_bounds__(64) void k_batch_restore_in(BatchArgs b, const int *nanb, const double *In_bak, double *In)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.nc || !nanb[b.camblk[k]]) return;
    for (int s = 0; s < 3; ++s) In[3 * (size_t)k + s] = In_bak[3 * (size_t)k + s];
}

__global__ __launch_bounds__(256) void k_copy_d(double *dst, const double *src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

void launch_block_sums(const BatchArgs &b, const double *terms, const float *wa, const float *wb,
                       const unsigned char *flag, double *out, double *part, hipStream_t st)
{
    k_block_sums<<<dim3(kSlices, b.nb), 256, 0, st>>>(b, terms, wa, wb, flag, part);
    k_block_sums_fin<<<nblk(b.nb, 64), 64, 0, st>>>(b.nb, part, out);
}
void launch_block_nan(const BatchArgs &b, const double *sol, const double *Xn, int *nanb, int *part, hipStream_t st)
{
    k_block_nan<<<dim3(kSlices, b.nb), 256, 0, st>>>(b, sol, Xn, part);
    k_block_nan_fin<<<nblk(b.nb, 64), 64, 0, st>>>(b.nb, part, nanb);
}
int block_slices() { return kSlices; }
void launch_batch_store(const BatchArgs &b, const unsigned char *accept, const CamPtrs &cp, const double *Xn,
                        double *X, const unsigned char *flag, unsigned char *flagst, hipStream_t st)
{
    const long long n = std::max(1, std::max(b.nc, b.npts));
    const int g = (int)std::min<long long>(nblk(n, 256), 4096);
    k_batch_store<<<g, 256, 0, st>>>(b, accept, cp, Xn, X, flag, flagst);
}
void launch_batch_restore_in(const BatchArgs &b, const int *nanb, const double *In_bak, double *In, hipStream_t st)
{
    k_batch_restore_in<<<nblk(std::max(b.nc, 1), 64), 64, 0, st>>>(b, nanb, In_bak, In);
}
// point -> block: the block whose union point range holds the point's union
// index (uptoff[lo] <= u < uptoff[lo + 1]; an empty block's equal offsets
// resolve to the block after it)
__global__ void k_batch_ptblk(int npts, int nb, const int *nrpv, const int *uptoff, int *ptblk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const int u = nrpv[i];
    int lo = 0, hi = nb;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (uptoff[mid] <= u)
            lo = mid;
        else
            hi = mid;
    }
    ptblk[i] = lo;
}

// block b's first kept point (the first with union index >= uptoff[b]; kept
// points stay in union order, so nrpv ascends) and first entry
__global__ void k_batch_ranges(int npts, int nb, const int *nrpv, const int *uptoff, const int *toff, int *poff,
                               int *eoff)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    const int key = uptoff[b];
    int lo = 0, hi = npts;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (nrpv[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    poff[b] = lo;
    eoff[b] = toff[lo];
}

void launch_batch_ranges(int npts, int nb, const int *nrpv, const int *uptoff, const int *toff, int *ptblk, int *poff,
                         int *eoff, hipStream_t st)
{
    if (npts > 0) k_batch_ptblk<<<nblk(npts, 256), 256, 0, st>>>(npts, nb, nrpv, uptoff, ptblk);
    k_batch_ranges<<<nblk(nb + 1, 64), 64, 0, st>>>(npts, nb, nrpv, uptoff, toff, poff, eoff);
}

void launch_copy(double *dst, const double *src, size_t n, hipStream_t st)
{
    if (!n) return;
    const int g = (int)std::min<size_t>((n + 255) / 256, 2048);
    k_copy_d<<<g, 256, 0, st>>>(dst, src, n);
}

}  // namespace rpba
