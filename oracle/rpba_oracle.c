/*
 * rpba_oracle.c -- CPU oracle for the global robust LM bundle adjustment of
 * helmayer/RPBA.  TEST INFRASTRUCTURE ONLY: the product (rpba_amd/, the C ABI
 * in include/rpba.h) never links, loads or calls this file.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * checker and as the timed CPU baseline.
 *
 * A from-scratch C restatement of the reference's algorithm (the reference is
 * C++ on Eigen 3, Boost and METIS, none of which exist in this image, so the
 * reference cannot be built here).  It keeps the reference's data
 * model: compact camera blocks of the sizes camparv gives (intrinsics only for
 * image 0, 8 for the gauge camera `normcamera`, 6+a for the others), the
 * camera-major observation numbering `nro`, the profile-style upper block
 * store of S with touched flags, and the reference's operation order in the
 * per-observation formulas.
 *
 * Two arithmetic modes:
 *   fsites = 1  the reference's float sites: the cost sig0 and the weight sum
 *               wsum accumulate in float in point order and the sigma algebra
 *               and LM comparisons run in float (rbacore.cpp:451, 487-494,
 *               526, 555-563, 727, 776, 881, 1206, 1312-1327);
 *   fsites = 0  plain FP64: the same sums in double, per-observation cost terms
 *               from the double residuals.  This is the GPU's plain form.
 * Per-observation float quantities that feed the robust labels (the key
 * sig0rmv, the weights Wij, the medians) are float in both modes.
 *
 * Parity pins (tests/test_oracle_*.py): Jacobians by central finite
 * differences; U/V/W/Schur/b by a dense numpy normal-equation build on tiny
 * scenes; PCG against a dense Cholesky solve and a textbook replay; Cayley,
 * rectification and angle-axis by closed forms; robust labels on constructed
 * outliers; the LM control flow by scripted cases.  Nothing pins the exact
 * Eigen summation order or GCC's FMA contraction of the reference build
 * (compiled here with -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/rpba.h"

#define MAXB 9

/* Test knobs, set only by tests/: a planted mistake (to show that the pins of
 * the LM control and of RobustDelete catch it) and a summation-order variant
 * (to measure how far rounding alone moves a trajectory).  Both are 0 in
 * every other use.
 *   mutation 1: after a rejected step the trial cameras are restored from the
 *               stored ones (H4 broken)
 *   mutation 2: the first iteration of a pass is stored only if better (H5)
 *   mutation 3: ISflagnew is copied from ISflag on every iteration (H6)
 *   mutation 4: RobustDelete drops a point with minimagenrdel left (<=)
 *   mutation 5: the itt cap breaks one iteration early (itt >= 200)
 *   mutation 6: the output intrinsics come from the stored state (H19)
 *   order 1:    S d sums its row products in the reverse column order
 *   order 2:    the Schur pair products are fused multiply-adds */
static int g_mut = 0;
static int g_order = 0;
void oracle_set_mutation(int m) { g_mut = m; }
void oracle_set_order(int m) { g_order = m; }

/* Arithmetic modes of the normal equations (the FP32 diagnostic
 * tests/golden/fp32_diagnostic.py, and the parity reference of the GPU's
 * whitened-Jacobian forms):
 *   prec 1: the GPU's FP32 records emulated -- the Jacobians Bc, Bp rounded
 *           to float, Bc^T W and W = Bc^T W Bp formed and rounded in float,
 *           Y = W V^-1 from a float V^-1 rounded in float; S sums the
 *           products of those floats in double;
 *   prec 2: consistently rounded factors -- with the weight W = L L^T, the
 *           whitened Jacobians L^T Bc and L^T Bp rounded to float once, and
 *           U, V, W, g, Y and S all formed from them in double;
 *   prec 3: the same whitened Jacobians, not rounded (FP64 throughout).
 * pcg_cap > 0 lowers the PCG cap; the step-norm ratios m / m0 of PCG solve
 * number pcg_trace_call (0-based, counted from the setting) go to
 * pcg_trace[0..cap); stop_after > 0 ends Adjust after that many LM
 * iterations. */
static int g_prec = 0, g_pcg_cap = 0, g_stop_after = 0;
static double *g_pcg_trace = NULL;
static int g_pcg_trace_cap = 0, g_pcg_trace_call = -1, g_pcg_calls = 0, g_pcg_trace_len = 0;
void oracle_set_diag(int prec, int pcg_cap, int stop_after, int trace_call, double *trace, int trace_cap)
{
    g_prec = prec;
    g_pcg_cap = pcg_cap;
    g_stop_after = stop_after;
    g_pcg_trace_call = trace_call;
    g_pcg_trace = trace;
    g_pcg_trace_cap = trace_cap;
    g_pcg_calls = 0;
    g_pcg_trace_len = 0;
}
int oracle_diag_trace_len(void) { return g_pcg_trace_len; }

/* ------------------------------------------------------------------------- */
/* small math                                                                */
/* ------------------------------------------------------------------------- */

/* Cayley rotation from (a,b,c); ComputeRotationMatrixfromRodrigues,
 * rbacore.cpp:2014-2043.  Row-major R. */
void oracle_cayley(double a, double b, double c, double *R)
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

/* Quaternion re-orthonormalisation of a rotation; RectifyRotationMatrix3/4,
 * rpba.h:876-960.  R row-major 3x3 with row stride `ld` (3 or 4). */
void oracle_rectify(double *R, int ld)
{
#define E(r, c) R[(r) * ld + (c)]
    double a0 = E(2, 1) - E(1, 2), a1 = E(0, 2) - E(2, 0), a2 = E(1, 0) - E(0, 1);
    if (fabs(a0) < 1.e-20) a0 = 1.e-20;
    if (fabs(a1) < 1.e-20) a1 = 1.e-20;
    if (fabs(a2) < 1.e-20) a2 = 1.e-20;
    const double al = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    a0 /= al; a1 /= al; a2 /= al;
    const double phi = atan2(al, E(0, 0) + E(1, 1) + E(2, 2) - 1.);
    const double q0 = cos(phi * 0.5), sh = sin(phi * 0.5);
    const double q1 = a0 * sh, q2 = a1 * sh, q3 = a2 * sh;
    E(0, 0) = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    E(0, 1) = 2 * (q1 * q2 - q0 * q3);
    E(0, 2) = 2 * (q1 * q3 + q0 * q2);
    E(1, 0) = 2 * (q2 * q1 + q0 * q3);
    E(1, 1) = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    E(1, 2) = 2 * (q2 * q3 - q0 * q1);
    E(2, 0) = 2 * (q3 * q1 - q0 * q2);
    E(2, 1) = 2 * (q3 * q2 + q0 * q1);
    E(2, 2) = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
#undef E
}

/* Angle-axis -> float rotation as InputBlock does (ComputeRotationMatrixfrom-
 * RodriguesAngle, rpba.cpp:343-366): double math, float result. */
void oracle_aa_to_R(double a, double b, double c, float *R)
{
    const double a2 = a * a, b2 = b * b, c2 = c * c;
    const double ab = a * b, ac = a * c, bc = b * c;
    const double aa = sqrt(a2 + b2 + c2);
    const double ct = aa == 0. ? 0.5 : (1. - cos(aa)) / (aa * aa);
    const double st = aa == 0. ? 1 : sin(aa) / aa;
    R[0] = (float)(1. - (b2 + c2) * ct);
    R[1] = (float)(ab * ct - c * st);
    R[2] = (float)(ac * ct + b * st);
    R[3] = (float)(ab * ct + c * st);
    R[4] = (float)(1. - (c2 + a2) * ct);
    R[5] = (float)(bc * ct - a * st);
    R[6] = (float)(ac * ct - b * st);
    R[7] = (float)(bc * ct + a * st);
    R[8] = (float)(1. - (a2 + b2) * ct);
}

/* Float rotation -> angle-axis as OutputBlock does
 * (ComputeRodriguesAnglefromRotationMatrix, rpba.cpp:371-439). */
void oracle_R_to_aa(const float *R, double *out)
{
    const double e0 = 0.01, e1 = 0.1, pi = 3.14159265358979323846;
#define E(r, c) ((double)R[(r) * 3 + (c)])
    const double tr = (E(0, 0) + E(1, 1) + E(2, 2) - 1.) * 0.5;
    if (fabs(E(0, 1) - E(1, 0)) < e0 && fabs(E(1, 2) - E(2, 1)) < e0 && fabs(E(0, 2) - E(2, 0)) < e0) {
        if (fabs(E(0, 1) + E(1, 0)) < e1 && fabs(E(1, 2) + E(2, 1)) < e1 && fabs(E(0, 2) + E(2, 0)) < e1 && tr > 0.9) {
            out[0] = out[1] = out[2] = 0.;
            return;
        }
        const double xx = (E(0, 0) + 1.) * 0.5, yy = (E(1, 1) + 1.) * 0.5, zz = (E(2, 2) + 1.) * 0.5;
        const double xy = (E(0, 1) + E(1, 0)) * 0.25, xz = (E(0, 2) + E(2, 0)) * 0.25, yz = (E(1, 2) + E(2, 1)) * 0.25;
        if (xx > yy && xx > zz) {
            if (xx < e0) { out[0] = 0.; out[1] = out[2] = sqrt(0.5) * pi; return; }
            const double t = sqrt(xx);
            out[0] = t * pi; out[1] = xy / t * pi; out[2] = xz / t * pi;
            return;
        }
        if (yy > zz) {
            if (yy < e0) { out[0] = out[2] = sqrt(0.5) * pi; out[1] = 0.; return; }
            const double t = sqrt(yy);
            out[0] = xy / t * pi; out[1] = t * pi; out[2] = yz / t * pi;
            return;
        }
        if (zz < e0) { out[0] = out[1] = sqrt(0.5) * pi; out[2] = 0.; return; }
        const double t = sqrt(zz);
        out[0] = xz / t * pi; out[1] = yz / t * pi; out[2] = t * pi;
        return;
    }
    const double aa = acos(tr);
    const double bb = 0.5 * aa / sin(aa);
    out[0] = bb * (E(2, 1) - E(1, 2));
    out[1] = bb * (E(0, 2) - E(2, 0));
    out[2] = bb * (E(1, 0) - E(0, 1));
#undef E
}

/* 3x3 inverse by cofactors (the reference calls Eigen's Matrix3d::inverse,
 * rbacore.cpp:1344, a cofactor/adjugate formula).  Row-major. */
void oracle_inv3(const double *m, double *r)
{
    /* cofactor c(i,j) = m(i1,j1) m(i2,j2) - m(i1,j2) m(i2,j1), i1=(i+1)%3 ... */
    double c[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            c[i * 3 + j] = m[i1 * 3 + j1] * m[i2 * 3 + j2] - m[i1 * 3 + j2] * m[i2 * 3 + j1];
        }
    const double det = c[0] * m[0] + c[3] * m[3] + c[6] * m[6];
    const double id = 1. / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r[i * 3 + j] = c[j * 3 + i] * id; /* adjugate = cofactor transpose */
}

/* n x n inverse by LU with partial (row) pivoting, the algorithm Eigen's
 * MatrixXd::inverse() uses for dynamic sizes (rbacore.cpp:1841).
 * Row-major, n <= MAXB. */
void oracle_invlu(int n, const double *a_in, double *inv)
{
    double a[MAXB * MAXB];
    int perm[MAXB];
    memcpy(a, a_in, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(a[i * n + k]) > best) { best = fabs(a[i * n + k]); p = i; }
        if (p != k) {
            for (int j = 0; j < n; ++j) { double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t; }
            int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
        }
        const double piv = a[k * n + k];
        if (piv != 0.)
            for (int i = k + 1; i < n; ++i) {
                const double l = a[i * n + k] / piv;
                a[i * n + k] = l;
                for (int j = k + 1; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
            }
    }
    /* solve L U X = P I column by column */
    for (int col = 0; col < n; ++col) {
        double y[MAXB];
        for (int i = 0; i < n; ++i) {
            double s = (perm[i] == col) ? 1. : 0.;
            for (int j = 0; j < i; ++j) s -= a[i * n + j] * y[j];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * y[j];
            y[i] = s / a[i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[i * n + col] = y[i];
    }
}

/* upper median of n floats: sorted[n/2] (rbacore.cpp:844-845) */
static int cmp_float(const void *a, const void *b)
{
    const float x = *(const float *)a, y = *(const float *)b;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}
float oracle_upper_median(const float *v, int n)
{
    float *t = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    memcpy(t, v, sizeof(float) * (size_t)n);
    qsort(t, (size_t)n, sizeof(float), cmp_float);
    const float m = t[n / 2];
    free(t);
    return m;
}

/* ------------------------------------------------------------------------- */
/* solver state                                                              */
/* ------------------------------------------------------------------------- */

/* per-LM-iteration records of the solver state (tests only): fingerprints
 * (position-weighted sums) of the camera and point states and robust keys */
typedef struct orc_rec {
    int cap, len;
    double init_key;           /* keys after the initial cost evaluation of pass 0 */
    double *lin_cam, *lin_pts; /* the state the linearisation (CalcBundleSep) used */
    double *trial_cam, *trial_pts, *trial_key; /* the trial after its decode and cost */
    double *stored_cam, *stored_pts;           /* the stored state after the decision */
    double *key_used;          /* keys Reweight2D labelled from (0: it did not run) */
    double *flag_fp, *flagnew_fp; /* fingerprints of ISflag / ISflagnew after the decision */
    int *iflag, *pass;
} orc_rec;

typedef struct orc {
    int fsites, robust, nraddpi, addf[3];
    int solver;              /* 0 = block-Jacobi PCG (the reference); 1 = dense Cholesky */
    double omega;            /* preconditioner 1: Neumann weight (0 = block-Jacobi) */
    int nimg, XXnr;
    float ihw2;
    double defdiff;
    const float *obs_xy, *obs_w;
    /* tracks (pointXXv after PointXXvIn) */
    int npts, ntr;
    int *nrpv, *toff, *timg, *tfo;
    float *wold;             /* Wijoldv, point-major [3*ntr] */
    /* camera-major numbering (nrobv / nrobvo) */
    int *nrobvo, *nro;
    float *Wij, *sig0rmv;
    unsigned char *ISflag, *ISflagnew;
    /* cameras: stored (PMatrices, cameras) and trial (eRMnew, ePCnew, camerasnew) */
    double *PM;              /* [12*nimg] row-major [R|t] */
    double *cf, *cd0, *cd1;
    double *nf, *nd0, *nd1;
    double *Rb, *Cb;         /* eRM / ePC, indexed by image (reference index k-1) */
    double *Rn, *Cn;         /* eRMnew / ePCnew */
    int normcamera, esvm[2], esvmaxm, lflflag, minimagenrdel;
    int nrpari, nrparir, nrioff, nrioff3;
    int *camparv;
    double *XP, *XPn, *XX, *XXn, *x;
    size_t xpsize;
    /* linear system */
    double **Sb;             /* [nimg*nimg] upper blocks k<=l, row-major bs_k x bs_l */
    unsigned char *Sf;
    double *ex;              /* exiv, flat [nrioff] */
    int begini;
    int maxtrack;
    /* scalars */
    double sig0, wsum, wsumg, sig0old, sig0in;
    int imagepointnr, nrobs, nrunknowns, redundancy, nrobsnew, nrunknownsnew, redundancynew;
    int64_t pcg_total;
    int pcg_last;
    int lasterr;
    /* tests only: scripted sig0 values replacing the computed ones, records */
    const double *script;
    int script_len, script_pos;
    orc_rec *rec;
    double *Rs, *Cs;         /* stored cameras in the (R, C) form of eRMnew / ePCnew */
    int pt_stride;           /* timing sample (orc_sample_iteration): every pt_stride-th point; 0 = all */
    /* GCP mode of a partition block (aoflag, rbacore.cpp:233-329, 1677-1729):
     * no gauge camera, every image with the full 6 + a block; gcpi[i] is the
     * kept point's index into the GCP centres gcpc (3 each) or -1; the n-th
     * kept GCP point takes entry n of the weight lists W3 (6 lists of nW3) */
    int aoflag, nrcam_unk;
    int *gcpi;
    const float *W3;
    int nW3;
    const double *gcpc;
    double sig0ao;
} orc;

#define FS(o, v) ((o)->fsites ? (double)(float)(v) : (double)(v))

static int bsz(const orc *o, int i) { return o->camparv[i + 1] - o->camparv[i]; }

static void orc_log(rpba_log_fn log, void *user, const char *s)
{
    if (log) log(s, user);
}

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    free(o->nrpv); free(o->toff); free(o->timg); free(o->tfo); free(o->wold);
    free(o->nrobvo); free(o->nro); free(o->Wij); free(o->sig0rmv); free(o->ISflag); free(o->ISflagnew);
    free(o->PM); free(o->cf); free(o->cd0); free(o->cd1); free(o->nf); free(o->nd0); free(o->nd1);
    free(o->Rb); free(o->Cb); free(o->Rn); free(o->Cn); free(o->camparv); free(o->Rs); free(o->Cs);
    free(o->XP); free(o->XPn); free(o->XX); free(o->XXn); free(o->x); free(o->ex);
    if (o->Sb) {
        for (size_t i = 0; i < (size_t)o->nimg * o->nimg; ++i) free(o->Sb[i]);
        free(o->Sb);
    }
    free(o->Sf);
    free(o->gcpi);
    free(o);
}

/* LSA::LSA + PointXXvIn + PrepareCameras + PrepareIteration
 * (rbacore.cpp:37-123, 1918-1974, 670-700, 127-229). */
orc *orc_create(const rpba_block *io, const rpba_flags *f, int fsites)
{
    orc *o = (orc *)xcalloc(1, sizeof(orc));
    const int nc = io->n_cams, np = io->n_pts, no = io->n_obs;
    o->fsites = fsites;
    o->solver = f->solver;
    o->omega = f->preconditioner == 1 ? (double)f->pc_omega : 0.;
    o->robust = f->robust != 0;
    o->addf[0] = f->est_k2 != 0; o->addf[1] = f->est_k4 != 0; o->addf[2] = f->est_f != 0;
    o->nraddpi = o->addf[0] + o->addf[1] + o->addf[2];
    o->nimg = nc;
    o->XXnr = np;
    o->obs_xy = io->obs_xy;
    o->obs_w = io->obs_w;
    /* ihw2 (rpbacore.cpp:87-90) */
    float ihw2 = 0.f;
    for (int i = 0; i < nc; ++i) {
        const int w = io->img_w ? io->img_w[i] : 1, h = io->img_h ? io->img_h[i] : 1;
        ihw2 += (float)((h + w) * (h + w)) * 0.25f;
    }
    ihw2 /= (float)nc;
    o->ihw2 = ihw2;
    o->defdiff = 5000. / sqrtf(ihw2); /* rbacore.cpp:48 */
    o->minimagenrdel = nc > 2 ? 3 : 2; /* rbacore.cpp:73-76 */

    /* tracks in file order (InputBlock rpba.cpp:219-238), then PointXXvIn */
    int *cnt = (int *)xcalloc((size_t)np + 1, sizeof(int));
    for (int i = 0; i < no; ++i)
        if (io->obs_cam[i] >= 0) cnt[io->obs_pt[i] + 1]++;
    for (int i = 0; i < np; ++i) cnt[i + 1] += cnt[i];
    int *all = (int *)xcalloc((size_t)cnt[np] + 1, sizeof(int));
    int *pos = (int *)xcalloc((size_t)np + 1, sizeof(int));
    memcpy(pos, cnt, sizeof(int) * (size_t)(np + 1));
    for (int i = 0; i < no; ++i)
        if (io->obs_cam[i] >= 0) all[pos[io->obs_pt[i]]++] = i;
    o->nrpv = (int *)xcalloc((size_t)np, sizeof(int));
    o->toff = (int *)xcalloc((size_t)np + 1, sizeof(int));
    o->timg = (int *)xcalloc((size_t)cnt[np] + 1, sizeof(int));
    o->tfo = (int *)xcalloc((size_t)cnt[np] + 1, sizeof(int));
    int npk = 0, nt = 0;
    for (int i = 0; i < np; ++i) {
        const int t = cnt[i + 1] - cnt[i];
        const double Z = (double)io->pts[3 * i + 2];
        if (fabs(Z) > 1.e-8 && t >= 2) { /* rbacore.cpp:1954-1955 */
            o->nrpv[npk] = i;
            o->toff[npk] = nt;
            for (int j = cnt[i]; j < cnt[i + 1]; ++j) {
                o->timg[nt] = io->obs_cam[all[j]];
                o->tfo[nt] = all[j];
                ++nt;
            }
            ++npk;
        }
    }
    o->toff[npk] = nt;
    o->npts = npk;
    o->ntr = nt;
    free(cnt); free(all); free(pos);

    o->XX = (double *)xcalloc((size_t)3 * npk + 3, sizeof(double));
    o->XXn = (double *)xcalloc((size_t)3 * npk + 3, sizeof(double));
    for (int i = 0; i < npk; ++i)
        for (int j = 0; j < 3; ++j)
            o->XX[3 * i + j] = o->XXn[3 * i + j] = (double)io->pts[3 * o->nrpv[i] + j];

    o->PM = (double *)xcalloc((size_t)12 * nc, sizeof(double));
    for (int i = 0; i < 12 * nc; ++i) o->PM[i] = (double)io->cam_P[i];

    /* PrepareCameras: gauge camera by |R_i t_i - R_0 t_0|^2 before any
     * rectification (rbacore.cpp:670-700) */
    {
        double T0[3];
        for (int r = 0; r < 3; ++r)
            T0[r] = o->PM[r * 4 + 0] * o->PM[3] + o->PM[r * 4 + 1] * o->PM[7] + o->PM[r * 4 + 2] * o->PM[11];
        double nd = 0.;
        o->normcamera = (nc > 1) ? 1 : -1; /* reference leaves it unset if no distance is > 0 */
        for (int i = 1; i < nc; ++i) {
            const double *P = o->PM + 12 * i;
            double d = 0.;
            for (int r = 0; r < 3; ++r) {
                const double T = P[r * 4 + 0] * P[3] + P[r * 4 + 1] * P[7] + P[r * 4 + 2] * P[11];
                d += (T - T0[r]) * (T - T0[r]);
            }
            if (d > nd) { nd = d; o->normcamera = i; }
        }
    }

    /* PrepareIteration (rbacore.cpp:127-229) */
    o->camparv = (int *)xcalloc((size_t)nc + 2, sizeof(int));
    o->Rb = (double *)xcalloc((size_t)9 * nc, sizeof(double));
    o->Cb = (double *)xcalloc((size_t)3 * nc, sizeof(double));
    o->Rn = (double *)xcalloc((size_t)9 * nc, sizeof(double));
    o->Cn = (double *)xcalloc((size_t)3 * nc, sizeof(double));
    o->Rs = (double *)xcalloc((size_t)9 * nc, sizeof(double));
    o->Cs = (double *)xcalloc((size_t)3 * nc, sizeof(double));
    oracle_rectify(o->PM, 4);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o->Rn[r * 3 + c] = o->PM[r * 4 + c];
    for (int c = 0; c < 3; ++c)
        o->Cn[c] = o->Rn[0 * 3 + c] * o->PM[3] + o->Rn[1 * 3 + c] * o->PM[7] + o->Rn[2 * 3 + c] * o->PM[11];
    o->nrparir = 5 + o->nraddpi;
    o->nrpari = 6 + o->nraddpi;
    o->cf = (double *)xcalloc((size_t)nc, sizeof(double));
    o->cd0 = (double *)xcalloc((size_t)nc, sizeof(double));
    o->cd1 = (double *)xcalloc((size_t)nc, sizeof(double));
    o->nf = (double *)xcalloc((size_t)nc, sizeof(double));
    o->nd0 = (double *)xcalloc((size_t)nc, sizeof(double));
    o->nd1 = (double *)xcalloc((size_t)nc, sizeof(double));
    o->lflflag = 0;
    for (int i = 0; i < nc; ++i) {
        o->cf[i] = o->nf[i] = (double)io->cam_f[i];
        o->cd0[i] = o->nd0[i] = (double)io->cam_d[2 * i];
        o->cd1[i] = o->nd1[i] = (double)io->cam_d[2 * i + 1];
        if (io->cam_f[i] > 2.f) o->lflflag = 1; /* rbacore.cpp:213-220 */
    }
    o->nrioff = o->nraddpi + o->nrparir + o->nrpari * (nc - 2);

    o->xpsize = (size_t)nc * 6 + (size_t)3 * np + (size_t)7 * nc + 16;
    o->XP = (double *)xcalloc(o->xpsize, sizeof(double));
    o->XPn = (double *)xcalloc(o->xpsize, sizeof(double));
    o->x = (double *)xcalloc(o->xpsize, sizeof(double));
    o->ex = (double *)xcalloc((size_t)o->nrioff + 16, sizeof(double));
    o->Sb = (double **)xcalloc((size_t)nc * nc, sizeof(double *));
    o->Sf = (unsigned char *)xcalloc((size_t)nc * nc, 1);
    o->nrobvo = (int *)xcalloc((size_t)nc + 2, sizeof(int));
    o->nro = (int *)xcalloc((size_t)o->ntr + 1, sizeof(int));
    o->Wij = (float *)xcalloc((size_t)3 * o->ntr + 3, sizeof(float));
    o->sig0rmv = (float *)xcalloc((size_t)o->ntr + 1, sizeof(float));
    o->ISflag = (unsigned char *)xcalloc((size_t)o->ntr + 1, 1);
    o->ISflagnew = (unsigned char *)xcalloc((size_t)o->ntr + 1, 1);
    o->wold = (float *)xcalloc((size_t)3 * o->ntr + 3, sizeof(float));
    for (int q = 0; q < o->ntr; ++q) {
        const int fo = o->tfo[q];
        o->wold[3 * q + 0] = o->obs_w ? o->obs_w[3 * fo + 0] : 1.f;
        o->wold[3 * q + 1] = o->obs_w ? o->obs_w[3 * fo + 1] : 1.f;
        o->wold[3 * q + 2] = o->obs_w ? o->obs_w[3 * fo + 2] : 0.f;
    }
    return o;
}

/* nrobv counts and the camera-major index of every track entry
 * (rbacore.cpp:401-430 and the running counters of 745-757). */
static int count_images(orc *o, rpba_log_fn log, void *user)
{
    const int nc = o->nimg;
    int *cnt = o->nrobvo;
    memset(cnt, 0, sizeof(int) * (size_t)(nc + 2));
    o->maxtrack = 0;
    for (int i = 0; i < o->npts; ++i) {
        const int t = o->toff[i + 1] - o->toff[i];
        if (t > o->maxtrack) o->maxtrack = t;
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q) cnt[o->timg[q] + 1]++;
    }
    for (int i = 1; i < nc + 1; ++i)
        if (cnt[i] == 0) {
            if (nc > 3) {
                char buf[128];
                snprintf(buf, sizeof buf, "BA-Error: No observations in image %d (image index)\n", i - 1);
                orc_log(log, user, buf);
            }
            return RPBA_E_NO_OBS_IN_IMAGE;
        }
    for (int i = 0; i < nc; ++i) cnt[i + 1] += cnt[i];
    int *run = (int *)xcalloc((size_t)nc + 1, sizeof(int));
    memcpy(run, cnt, sizeof(int) * (size_t)nc);
    for (int q = 0; q < o->ntr; ++q) o->nro[q] = run[o->timg[q]]++;
    free(run);
    return 0;
}

/* PrepareWeights2D (rbacore.cpp:704-783). Wij in camera-major order. */
static void prepare_weights(orc *o)
{
    double ws = 0.;
    for (int q = 0; q < o->ntr; ++q) {
        const float *w = o->wold + 3 * q; /* input weights (Wijoldv or wx/wy/wxy) */
        float *W = o->Wij + 3 * o->nro[q];
        W[0] = w[0]; W[1] = w[1]; W[2] = w[2];
        ws = o->fsites ? (double)((float)ws + (w[0] + w[1])) : ws + ((double)w[0] + (double)w[1]);
    }
    o->wsum = ws;
}

/* EncodeCameras (rbacore.cpp:999-1087), non-GCP branch. */
static void encode_cameras(orc *o)
{
    const int nc = o->nimg;
    for (int k = 0; k < nc + 1; ++k) o->camparv[k] = 0;
    if (o->aoflag) {
        /* GCP branch (rbacore.cpp:1005-1009, 1025-1082): every image, image 0
         * included, is re-orthonormalised and gets a block of nrpari slots
         * (pose increments zero, then its own intrinsics) */
        for (int k = 0; k < nc; ++k) {
            const int cp = o->camparv[k];
            double *P = o->PM + 12 * k;
            oracle_rectify(P, 4);
            double *R = o->Rb + 9 * k, *C = o->Cb + 3 * k;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) R[r * 3 + c] = P[r * 4 + c];
            for (int c = 0; c < 3; ++c) C[c] = R[0 * 3 + c] * P[3] + R[1 * 3 + c] * P[7] + R[2 * 3 + c] * P[11];
            for (int i = 0; i < 6; ++i) o->XP[cp + i] = 0.;
            int a = 5 + cp;
            if (o->addf[0]) o->XP[++a] = o->cd0[k];
            if (o->addf[1]) o->XP[++a] = o->cd1[k];
            if (o->addf[2]) o->XP[++a] = o->cf[k];
            o->camparv[k + 1] = cp + o->nrpari;
        }
        o->nrioff3 = o->nrioff + 3 * o->npts;
        return;
    }
    o->camparv[1] = o->nraddpi;
    int addp = -1;
    if (o->addf[0]) o->XP[++addp] = o->cd0[0];
    if (o->addf[1]) o->XP[++addp] = o->cd1[0];
    if (o->addf[2]) o->XP[++addp] = o->cf[0];
    for (int k = 1; k < nc; ++k) {
        const int cp = o->camparv[k];
        double *P = o->PM + 12 * k;
        oracle_rectify(P, 4);
        double *R = o->Rb + 9 * k, *C = o->Cb + 3 * k;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) R[r * 3 + c] = P[r * 4 + c];
        for (int c = 0; c < 3; ++c) C[c] = R[0 * 3 + c] * P[3] + R[1 * 3 + c] * P[7] + R[2 * 3 + c] * P[11];
        for (int i = 0; i < 3; ++i) o->XP[cp + i] = 0.;
        if (k == o->normcamera) {
            float maxv = 0.f;
            int maxi = 0;
            for (int i = 0; i < 3; ++i)
                if (fabs(C[i]) > maxv) { maxv = (float)fabs(C[i]); maxi = i; }
            int ii = -1;
            for (int i = 0; i < 3; ++i)
                if (maxi != i) o->esvm[++ii] = i;
            o->esvmaxm = maxi;
            o->XP[cp + 3] = o->XP[cp + 4] = 0.;
            int a = 4 + cp;
            if (o->addf[0]) o->XP[++a] = o->cd0[k];
            if (o->addf[1]) o->XP[++a] = o->cd1[k];
            if (o->addf[2]) o->XP[++a] = o->cf[k];
            o->camparv[k + 1] = cp + o->nrparir;
        } else {
            for (int i = 0; i < 3; ++i) o->XP[cp + 3 + i] = 0.;
            int a = 5 + cp;
            if (o->addf[0]) o->XP[++a] = o->cd0[k];
            if (o->addf[1]) o->XP[++a] = o->cd1[k];
            if (o->addf[2]) o->XP[++a] = o->cf[k];
            o->camparv[k + 1] = cp + o->nrpari;
        }
    }
    o->nrioff3 = o->nrioff + 3 * o->npts;
}

/* DecodeCameras (rbacore.cpp:1733-1800), non-GCP branch: image 0 keeps its
 * pose; every other image gets Cay(theta_rot) * base, rectified. */
static void decode_cameras(orc *o, const double *XM)
{
    /* GCP mode: image 0 is decoded like the others (rbacore.cpp:1740) */
    for (int k = o->aoflag ? 0 : 1; k < o->nimg; ++k) {
        const int cp = o->camparv[k];
        double Ra[9], *Rn = o->Rn + 9 * k;
        const double *Rb = o->Rb + 9 * k;
        oracle_cayley(XM[cp], XM[cp + 1], XM[cp + 2], Ra);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                Rn[r * 3 + c] = Ra[r * 3 + 0] * Rb[0 * 3 + c] + Ra[r * 3 + 1] * Rb[1 * 3 + c] + Ra[r * 3 + 2] * Rb[2 * 3 + c];
        oracle_rectify(Rn, 3);
        double e[3];
        int a;
        if (k == o->normcamera) {
            e[o->esvm[0]] = XM[cp + 3];
            e[o->esvm[1]] = XM[cp + 4];
            e[o->esvmaxm] = 0.;
            a = 4;
        } else {
            for (int i = 0; i < 3; ++i) e[i] = XM[cp + 3 + i];
            a = 5;
        }
        for (int i = 0; i < 3; ++i) o->Cn[3 * k + i] = e[i] + o->Cb[3 * k + i];
        if (o->addf[0]) o->nd0[k] = XM[++a + cp];
        if (o->addf[1]) o->nd1[k] = XM[++a + cp];
        if (o->addf[2]) o->nf[k] = XM[++a + cp];
    }
    if (o->aoflag) return;
    int a = -1;
    if (o->addf[0]) o->nd0[0] = XM[++a];
    if (o->addf[1]) o->nd1[0] = XM[++a];
    if (o->addf[2]) o->nf[0] = XM[++a];
}

/* One observation's projection and residual at the trial cameras and the point
 * X (CalcCorrections rbacore.cpp:1238-1307, RadialReconstructionImage
 * rpba.h:850-857).  Returns the nan flag. */
typedef struct {
    double u, v;       /* ex3(0), ex3(1) after the 1/z scaling */
    double xs3;        /* guarded 1/e_z */
    double su, sv;     /* ex3u: distorted normalised coordinates */
    double diff1, diff2;
    int nan;
} obs_eval_t;

static void eval_obs(const orc *o, int k, const double *X, const float *xy, obs_eval_t *e)
{
    const double *R = o->Rn + 9 * k, *C = o->Cn + 3 * k;
    const double p0 = X[0] + C[0], p1 = X[1] + C[1], p2 = X[2] + C[2];
    double ex0 = R[0] * p0 + R[1] * p1 + R[2] * p2;
    double ex1 = R[3] * p0 + R[4] * p1 + R[5] * p2;
    double xs3 = R[6] * p0 + R[7] * p1 + R[8] * p2;
    if (fabs(xs3) < 1.e-3)
        xs3 = (xs3 < 0.) ? -1.e3 : 1.e3;
    else
        xs3 = 1. / xs3;
    ex0 *= xs3;
    ex1 *= xs3;
    e->u = ex0; e->v = ex1; e->xs3 = xs3;
    const double dist = ex0 * ex0 + ex1 * ex1;
    const double z = 1. + (o->nd0[k] * dist + o->nd1[k] * dist * dist);
    e->su = ex0 * z;
    e->sv = ex1 * z;
    const double f = o->nf[k];
    double d1 = (double)xy[0] - e->su * f;
    double d2 = (double)xy[1] - e->sv * f;
    if (d1 > o->defdiff) d1 = o->defdiff;
    if (d1 < -o->defdiff) d1 = -o->defdiff;
    if (d2 > o->defdiff) d2 = o->defdiff;
    if (d2 < -o->defdiff) d2 = -o->defdiff;
    e->nan = isnan(d1) || isinf(d1) || isnan(d2) || isinf(d2);
    if (e->nan) d1 = d2 = 0.;
    e->diff1 = d1; e->diff2 = d2;
}

/* Calibrated derivatives (CalcDerivatives rbacore.cpp:1420-1616 with
 * eCalcCalibratedDerivativesX/3, 1620-1673).  Fills Bc (2 x n, row-major with
 * row stride MAXB; n = the camera's compact block size) and Bp (2x3). */
static int derivs(const orc *o, int k, const obs_eval_t *e, double *Bc, double *Bp)
{
    const double *R = o->Rn + 9 * k;
    const double u = e->u, v = e->v, xs3 = e->xs3;
    const double x1 = u * xs3, x2 = v * xs3;
    const double xs31 = 1. / xs3, xs32 = xs31 * xs31;
    const double k2 = o->nd0[k], k4 = o->nd1[k], f = o->nf[k];
    const double xy2 = u * u + v * v, xy4 = xy2 * xy2;
    const double xxpc = k4 * xy4 + u * (2 * k2 * u + 4 * k4 * u * xy2) + k2 * xy2 + 1.;
    const double xypc = u * (4 * k4 * xy2 * v + 2 * k2 * v);
    const double yxpc = (4 * k4 * u * xy2 + 2 * k2 * u) * v;
    const double yypc = v * (2 * k2 * v + 4 * k4 * xy2 * v) + k4 * xy4 + k2 * xy2 + 1.;
#define CAL(c0, c1, o0, o1)                                  \
    do {                                                     \
        const double xp_ = xxpc * (c0) + xypc * (c1);        \
        const double yp_ = yypc * (c1) + yxpc * (c0);        \
        (o0) = xp_ * f + yp_ * 0.;                           \
        (o1) = yp_ * f;                                      \
    } while (0)
    /* point columns: (-R0j xs3 + R2j x1, -R1j xs3 + R2j x2), calibrated */
    for (int j = 0; j < 3; ++j) {
        const double c0 = -R[0 * 3 + j] * xs3 + R[2 * 3 + j] * x1;
        const double c1 = -R[1 * 3 + j] * xs3 + R[2 * 3 + j] * x2;
        CAL(c0, c1, Bp[j], Bp[3 + j]);
    }
    const double xyp2 = u * u + v * v, xyp4 = xyp2 * xyp2;
    double id0[2], id1[2], iff[2];
    id0[0] = -u * xyp2 * f - v * xyp2 * 0.; id0[1] = -v * xyp2 * f;
    id1[0] = -u * xyp4 * f - v * xyp4 * 0.; id1[1] = -v * xyp4 * f;
    iff[0] = -e->su; iff[1] = -e->sv;
    int n = 0;
    if (k > 0 || o->aoflag) { /* k >= begini (rbacore.cpp:1441) */
        double rc[2][3];
        rc[0][0] = x2 * x1 * xs32;       rc[1][0] = 1. + x2 * x2 * xs32;
        rc[0][1] = -1. - x1 * x1 * xs32; rc[1][1] = -x2 * x1 * xs32;
        rc[0][2] = x2 * xs31;            rc[1][2] = -x1 * xs31;
        for (int j = 0; j < 3; ++j) { CAL(rc[0][j], rc[1][j], Bc[j], Bc[MAXB + j]); }
        n = 3;
        if (k == o->normcamera) {
            for (int jj = 0; jj < 2; ++jj) { Bc[n] = Bp[o->esvm[jj]]; Bc[MAXB + n] = Bp[3 + o->esvm[jj]]; ++n; }
        } else {
            for (int j = 0; j < 3; ++j) { Bc[n] = Bp[j]; Bc[MAXB + n] = Bp[3 + j]; ++n; }
        }
    }
    if (o->addf[0]) { Bc[n] = id0[0]; Bc[MAXB + n] = id0[1]; ++n; }
    if (o->addf[1]) { Bc[n] = id1[0]; Bc[MAXB + n] = id1[1]; ++n; }
    if (o->addf[2]) { Bc[n] = iff[0]; Bc[MAXB + n] = iff[1]; ++n; }
#undef CAL
    return n;
}

/* S block (k,l), k<=l, allocated on first touch like the profile store of
 * rbacore.cpp:1364-1389 */
static double *sblock(orc *o, int k, int l)
{
    const size_t id = (size_t)k * o->nimg + l;
    if (!o->Sb[id]) o->Sb[id] = (double *)xcalloc((size_t)MAXB * MAXB, sizeof(double));
    o->Sf[id] = 1;
    return o->Sb[id];
}

/* CalcCorrections (rbacore.cpp:1201-1414).  mode: 0 = cost only,
 * 1 = linearise + Schur, 2 = back-substitution into x. */
static void corrections(orc *o, int mode, double mult1, const double *XM, int srmv)
{
    double sig0 = 0., sao = 0.;
    int gcpnum = 0;
    int kkk = o->nrioff - 1;
    const int mt = o->maxtrack > 0 ? o->maxtrack : 1;
    double *Wl = (double *)xcalloc((size_t)mt * MAXB * 3, sizeof(double)); /* eNlij */
    int *nl = (int *)xcalloc((size_t)mt, sizeof(int));
    const int pstride = o->pt_stride > 1 ? o->pt_stride : 1;
    for (int i = 0; i < o->npts; i += pstride) {
        double V[9] = {0}, g[3] = {0};
        double *X = o->XXn + 3 * i;
        for (int j = 0; j < 3; ++j) X[j] = XM[o->nrioff + 3 * i + j];
        int ks = 0;
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q, ++ks) {
            const int k = o->timg[q], nro = o->nro[q], fo = o->tfo[q];
            obs_eval_t e;
            eval_obs(o, k, X, o->obs_xy + 2 * fo, &e);
            const float *W = o->Wij + 3 * nro;
            if (mode > 0) {
                double Bc[2 * MAXB], Bp[6];
                const int n = derivs(o, k, &e, Bc, Bp);
                nl[ks] = n;
                double Bct[MAXB][2], Bpt[3][2];
                float Wunit[3] = {1.f, 1.f, 0.f};
                if (g_prec == 1) {
                    for (int j = 0; j < n; ++j) { Bc[j] = (float)Bc[j]; Bc[MAXB + j] = (float)Bc[MAXB + j]; }
                    for (int j = 0; j < 3; ++j) { Bp[j] = (float)Bp[j]; Bp[3 + j] = (float)Bp[3 + j]; }
                } else if (g_prec == 2 || g_prec == 3) {
                    /* W = L L^T; whitened columns L^T B (rounded once to
                     * float in mode 2); residual L^T r */
                    const double l11 = sqrt((double)W[0]), l21 = (double)W[2] / l11;
                    const double l22 = sqrt((double)W[1] - l21 * l21);
                    for (int j = 0; j < n; ++j) {
                        const double a0 = l11 * Bc[j] + l21 * Bc[MAXB + j], a1 = l22 * Bc[MAXB + j];
                        Bc[j] = g_prec == 2 ? (double)(float)a0 : a0;
                        Bc[MAXB + j] = g_prec == 2 ? (double)(float)a1 : a1;
                    }
                    for (int j = 0; j < 3; ++j) {
                        const double a0 = l11 * Bp[j] + l21 * Bp[3 + j], a1 = l22 * Bp[3 + j];
                        Bp[j] = g_prec == 2 ? (double)(float)a0 : a0;
                        Bp[3 + j] = g_prec == 2 ? (double)(float)a1 : a1;
                    }
                    const double r0 = l11 * e.diff1 + l21 * e.diff2, r1 = l22 * e.diff2;
                    e.diff1 = r0;
                    e.diff2 = r1;
                    W = Wunit;
                }
                for (int j = 0; j < n; ++j) {
                    const double xp = Bc[j], yp = Bc[MAXB + j];
                    Bct[j][0] = xp * W[0] + yp * W[2];
                    Bct[j][1] = xp * W[2] + yp * W[1];
                    if (g_prec == 1) { Bct[j][0] = (float)Bct[j][0]; Bct[j][1] = (float)Bct[j][1]; }
                }
                if (mode == 1 && n > 0) {
                    double *U = sblock(o, k, k);
                    double *gc = o->ex + o->camparv[k];
                    for (int r = 0; r < n; ++r) {
                        for (int c = 0; c < n; ++c) U[r * MAXB + c] += Bct[r][0] * Bc[c] + Bct[r][1] * Bc[MAXB + c];
                        gc[r] += Bct[r][0] * e.diff1 + Bct[r][1] * e.diff2;
                    }
                }
                double *Wq = Wl + (size_t)ks * MAXB * 3;
                for (int r = 0; r < n; ++r)
                    for (int c = 0; c < 3; ++c) {
                        Wq[r * 3 + c] = Bct[r][0] * Bp[c] + Bct[r][1] * Bp[3 + c];
                        if (g_prec == 1) Wq[r * 3 + c] = (float)Wq[r * 3 + c];
                    }
                for (int j = 0; j < 3; ++j) {
                    const double xp = Bp[j], yp = Bp[3 + j];
                    Bpt[j][0] = xp * W[0] + yp * W[2];
                    Bpt[j][1] = xp * W[2] + yp * W[1];
                }
                for (int r = 0; r < 3; ++r) {
                    for (int c = 0; c < 3; ++c) V[r * 3 + c] += Bpt[r][0] * Bp[c] + Bpt[r][1] * Bp[3 + c];
                    g[r] += Bpt[r][0] * e.diff1 + Bpt[r][1] * e.diff2;
                }
            }
            if (e.nan) {
                o->sig0rmv[nro] = 1.e20f;
            } else {
                const float d1f = (float)e.diff1, d2f = (float)e.diff2;
                const float d11 = d1f * d1f, d12 = 2.f * d1f * d2f, d22 = d2f * d2f;
                const float sv = d11 * W[0] + d12 * W[2] + d22 * W[1];
                float srv = sv;
                if (o->robust) {
                    const float *w0 = o->wold + 3 * q;
                    srv = d11 * w0[0] + d12 * w0[2] + d22 * w0[1];
                }
                if (srmv) o->sig0rmv[nro] = srv;
                if (o->fsites)
                    sig0 = (double)((float)sig0 + sv);
                else
                    sig0 += e.diff1 * e.diff1 * (double)W[0] + 2. * e.diff1 * e.diff2 * (double)W[2] +
                            e.diff2 * e.diff2 * (double)W[1];
            }
        }
        if (o->aoflag && o->gcpi[i] > -1) {
            /* CalcDerivatives3DGCP (rbacore.cpp:1677-1729): the prior on the
             * point's 3D position, B = -I, weight W3 of this GCP number */
            const int gn = gcpnum++;
            const double *cg = o->gcpc + 3 * (size_t)o->gcpi[i];
            const double d0 = cg[0] - X[0], d1 = cg[1] - X[1], d2 = cg[2] - X[2];
            double w[6];
            for (int v = 0; v < 6; ++v) w[v] = (double)o->W3[(size_t)v * o->nW3 + gn];
            sao += d0 * d0 * w[0] + d1 * d1 * w[1] + d2 * d2 * w[2] + 2 * (d0 * d1 * w[3] + d0 * d2 * w[4] + d1 * d2 * w[5]);
            if (mode > 0) {
                const double Ws[3][3] = {{w[0], w[3], w[4]}, {w[3], w[1], w[5]}, {w[4], w[5], w[2]}};
                double Bt[3][3];
                for (int j = 0; j < 3; ++j)
                    for (int c = 0; c < 3; ++c) Bt[j][c] = -Ws[j][c];
                const double dd[3] = {d0, d1, d2};
                for (int j = 0; j < 3; ++j) {
                    for (int c = 0; c < 3; ++c) V[j * 3 + c] += -Bt[j][c];
                    g[j] += Bt[j][0] * dd[0] + Bt[j][1] * dd[1] + Bt[j][2] * dd[2];
                }
            }
        }
        if (mode == 0) continue;
        /* V damping and inverse (rbacore.cpp:1341-1348) */
        for (int j = 0; j < 3; ++j) V[j * 3 + j] *= mult1;
        double Vi[9];
        oracle_inv3(V, Vi);
        const int t = o->toff[i + 1] - o->toff[i];
        if (mode == 1) {
            /* Schur pairs (rbacore.cpp:1350-1392) */
            for (int is = 0; is < t; ++is) {
                const int ij = o->timg[o->toff[i] + is];
                const int ni = nl[is];
                if (ni == 0) continue; /* image 0 without intrinsics: ij < begini */
                const double *Wa = Wl + (size_t)is * MAXB * 3;
                double N[MAXB * 3]; /* eNlijN = eNlij * V^-1 */
                for (int r = 0; r < ni; ++r)
                    for (int c = 0; c < 3; ++c) {
                        if (g_prec == 1) {
                            const float v0 = (float)Vi[0 * 3 + c], v1 = (float)Vi[1 * 3 + c], v2 = (float)Vi[2 * 3 + c];
                            const float w0 = (float)Wa[r * 3 + 0], w1 = (float)Wa[r * 3 + 1], w2 = (float)Wa[r * 3 + 2];
                            N[r * 3 + c] = (float)((float)((float)(w0 * v0) + (float)(w1 * v1)) + (float)(w2 * v2));
                        } else
                            N[r * 3 + c] = Wa[r * 3 + 0] * Vi[0 * 3 + c] + Wa[r * 3 + 1] * Vi[1 * 3 + c] + Wa[r * 3 + 2] * Vi[2 * 3 + c];
                    }
                double *D = sblock(o, ij, ij);
                double *gc = o->ex + o->camparv[ij];
                for (int r = 0; r < ni; ++r) {
                    for (int c = 0; c < ni; ++c) {
                        if (g_order & 2)
                            D[r * MAXB + c] = fma(-N[r * 3 + 2], Wa[c * 3 + 2],
                                                  fma(-N[r * 3 + 1], Wa[c * 3 + 1],
                                                      fma(-N[r * 3 + 0], Wa[c * 3 + 0], D[r * MAXB + c])));
                        else
                            D[r * MAXB + c] -= N[r * 3 + 0] * Wa[c * 3 + 0] + N[r * 3 + 1] * Wa[c * 3 + 1] + N[r * 3 + 2] * Wa[c * 3 + 2];
                    }
                    gc[r] -= N[r * 3 + 0] * g[0] + N[r * 3 + 1] * g[1] + N[r * 3 + 2] * g[2];
                }
                for (int iis = is + 1; iis < t; ++iis) {
                    const int ii = o->timg[o->toff[i] + iis];
                    const int nb = nl[iis];
                    if (nb == 0) continue;
                    const double *Wb = Wl + (size_t)iis * MAXB * 3;
                    if (ij < ii) {
                        double *B = sblock(o, ij, ii);
                        for (int r = 0; r < ni; ++r)
                            for (int c = 0; c < nb; ++c) {
                                if (g_order & 2)
                                    B[r * MAXB + c] = fma(-N[r * 3 + 2], Wb[c * 3 + 2],
                                                          fma(-N[r * 3 + 1], Wb[c * 3 + 1],
                                                              fma(-N[r * 3 + 0], Wb[c * 3 + 0], B[r * MAXB + c])));
                                else
                                    B[r * MAXB + c] -= N[r * 3 + 0] * Wb[c * 3 + 0] + N[r * 3 + 1] * Wb[c * 3 + 1] + N[r * 3 + 2] * Wb[c * 3 + 2];
                            }
                    } else {
                        /* ij > ii: block (ii, ij) -= eNlij[iis] * eNlijN^T; ij == ii
                         * cannot occur (duplicate cameras in a track are rejected) */
                        double *B = sblock(o, ii, ij);
                        for (int r = 0; r < nb; ++r)
                            for (int c = 0; c < ni; ++c) {
                                if (g_order & 2)
                                    B[r * MAXB + c] = fma(-Wb[r * 3 + 2], N[c * 3 + 2],
                                                          fma(-Wb[r * 3 + 1], N[c * 3 + 1],
                                                              fma(-Wb[r * 3 + 0], N[c * 3 + 0], B[r * MAXB + c])));
                                else
                                    B[r * MAXB + c] -= Wb[r * 3 + 0] * N[c * 3 + 0] + Wb[r * 3 + 1] * N[c * 3 + 1] + Wb[r * 3 + 2] * N[c * 3 + 2];
                            }
                    }
                }
            }
        } else {
            /* back-substitution (rbacore.cpp:1395-1410) */
            double xp[3];
            for (int r = 0; r < 3; ++r) xp[r] = Vi[r * 3 + 0] * g[0] + Vi[r * 3 + 1] * g[1] + Vi[r * 3 + 2] * g[2];
            for (int is = 0; is < t; ++is) {
                const int j = o->timg[o->toff[i] + is];
                const int n = nl[is];
                if (n == 0) continue;
                const double *Wa = Wl + (size_t)is * MAXB * 3;
                const double *xc = o->ex + o->camparv[j];
                double T[3 * MAXB]; /* V^-1 * eNlij^T */
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < n; ++c)
                        T[r * MAXB + c] = Vi[r * 3 + 0] * Wa[c * 3 + 0] + Vi[r * 3 + 1] * Wa[c * 3 + 1] + Vi[r * 3 + 2] * Wa[c * 3 + 2];
                for (int r = 0; r < 3; ++r) {
                    double s = 0.;
                    for (int c = 0; c < n; ++c) s += T[r * MAXB + c] * xc[c];
                    xp[r] -= s;
                }
            }
            o->x[++kkk] = xp[0];
            o->x[++kkk] = xp[1];
            o->x[++kkk] = xp[2];
        }
    }
    free(Wl);
    free(nl);
    o->sig0 = sig0;
    o->sig0ao = sao;
}

/* y = S d over the touched blocks, both triangles (eAAdv rbacore.cpp:2048-2068) */
static void spmv(int nb, int begini, const int *off, double *const *Sb, const unsigned char *Sf,
                 const double *d, double *y)
{
    for (int i = begini; i < nb; ++i)
        for (int j = off[i]; j < off[i + 1]; ++j) y[j] = 0.;
    for (int l = begini; l < nb; ++l) {
        const int nl_ = off[l + 1] - off[l];
        for (int k = l; k >= begini; --k) {
            const size_t id = (size_t)k * nb + l;
            if (!Sf[id]) continue;
            const double *B = Sb[id];
            const int nk = off[k + 1] - off[k];
            for (int r = 0; r < nk; ++r) {
                double s = 0.;
                if (g_order & 1)
                    for (int c = nl_ - 1; c >= 0; --c) s += B[r * MAXB + c] * d[off[l] + c];
                else
                    for (int c = 0; c < nl_; ++c) s += B[r * MAXB + c] * d[off[l] + c];
                y[off[k] + r] += s;
            }
            if (k != l)
                for (int c = 0; c < nl_; ++c) {
                    double s = 0.;
                    for (int r = 0; r < nk; ++r) s += B[r * MAXB + c] * d[off[k] + r];
                    y[off[l] + c] += s;
                }
        }
    }
}

/* sum over blocks of the block dot products (evvdv rbacore.cpp:2072-2084) */
static double dotb(int nb, int begini, const int *off, const double *a, const double *b)
{
    double c = 0.;
    for (int i = begini; i < nb; ++i) {
        double cc = 0.;
        for (int j = off[i]; j < off[i + 1]; ++j) cc += b[j] * a[j];
        c += cc;
    }
    return c;
}

/* y = N d over the off-diagonal touched blocks (both triangles) */
static void spmv_off(int nb, int begini, const int *off, double *const *Sb, const unsigned char *Sf,
                     const double *d, double *y)
{
    for (int i = begini; i < nb; ++i)
        for (int j = off[i]; j < off[i + 1]; ++j) y[j] = 0.;
    for (int l = begini; l < nb; ++l) {
        const int nl_ = off[l + 1] - off[l];
        for (int k = l - 1; k >= begini; --k) {
            const size_t id = (size_t)k * nb + l;
            if (!Sf[id]) continue;
            const double *B = Sb[id];
            const int nk = off[k + 1] - off[k];
            for (int r = 0; r < nk; ++r) {
                double s = 0.;
                for (int c = 0; c < nl_; ++c) s += B[r * MAXB + c] * d[off[l] + c];
                y[off[k] + r] += s;
            }
            for (int c = 0; c < nl_; ++c) {
                double s = 0.;
                for (int r = 0; r < nk; ++r) s += B[r * MAXB + c] * d[off[k] + r];
                y[off[l] + c] += s;
            }
        }
    }
}

/* preconditioner: h = C r (block-Jacobi), or with omega > 0 the first-order
 * Neumann correction h = C r - omega C N C r (N = off-diagonal part of S) */
static void precond(int nb, int begini, const int *off, double *const *Sb, const unsigned char *Sf,
                    const double *Cinv, double omega, const double *r, double *h, double *t, double *u)
{
    for (int i = begini; i < nb; ++i) {
        const int m = off[i + 1] - off[i];
        const double *Ci = Cinv + (size_t)i * MAXB * MAXB;
        for (int rr = 0; rr < m; ++rr) {
            double s = 0.;
            for (int cc = 0; cc < m; ++cc) s += Ci[rr * m + cc] * r[off[i] + cc];
            h[off[i] + rr] = s;
        }
    }
    if (omega == 0.) return;
    for (int j = off[begini]; j < off[nb]; ++j) t[j] = h[j];
    spmv_off(nb, begini, off, Sb, Sf, t, u);
    for (int i = begini; i < nb; ++i) {
        const int m = off[i + 1] - off[i];
        const double *Ci = Cinv + (size_t)i * MAXB * MAXB;
        for (int rr = 0; rr < m; ++rr) {
            double s = 0.;
            for (int cc = 0; cc < m; ++cc) s += Ci[rr * m + cc] * u[off[i] + cc];
            h[off[i] + rr] = t[off[i] + rr] - omega * s;
        }
    }
}

/* Block-Jacobi PCG with the reference's step-norm stopping rule and
 * best-iterate return (SparsePCG rbacore.cpp:1807-1914).  b in, x out.
 * omega > 0 adds the first-order Neumann correction to the preconditioner
 * (the corrected form; the reference has omega = 0).  Returns the iteration
 * count. */
int oracle_block_pcg_w(int nb, int begini, const int *off, double *const *Sb, const unsigned char *Sf,
                       const double *b, double *xout, double omega)
{
    const int n = off[nb];
    double *x = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *xo = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *r = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *r1 = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *d = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *Ad = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *h = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *h1 = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *hd = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *tw = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *uw = (double *)xcalloc((size_t)n + 1, sizeof(double));
    double *Cinv = (double *)xcalloc((size_t)nb * MAXB * MAXB, sizeof(double));
    /* every vector starts as a copy of b; x = 0 (rbacore.cpp:1826-1836) */
    for (int j = 0; j < n; ++j) r[j] = r1[j] = d[j] = Ad[j] = h[j] = h1[j] = xo[j] = b[j];
    int msize = 0;
    for (int i = begini; i < nb; ++i) {
        const int m = off[i + 1] - off[i];
        double a[MAXB * MAXB];
        for (int rr = 0; rr < m; ++rr)
            for (int cc = 0; cc < m; ++cc) a[rr * m + cc] = Sb[(size_t)i * nb + i][rr * MAXB + cc];
        if (m > 0) oracle_invlu(m, a, Cinv + (size_t)i * MAXB * MAXB);
        msize += m;
    }
    precond(nb, begini, off, Sb, Sf, Cinv, omega, r, h, tw, uw);
    for (int j = off[begini]; j < off[nb]; ++j) d[j] = h[j];
    if (msize < 1000) msize = 1000;
    if (g_pcg_cap > 0 && msize > g_pcg_cap) msize = g_pcg_cap;
    const int trace_this = (g_pcg_trace && g_pcg_calls == g_pcg_trace_call);
    if (g_pcg_trace) ++g_pcg_calls; /* only while a diagnostic traces (the partitioned solve runs blocks in threads) */
    double mold = 1.e50, m0 = 1.e50;
    int k;
    for (k = 0; k < msize; ++k) {
        spmv(nb, begini, off, Sb, Sf, d, Ad);
        const double rth = dotb(nb, begini, off, r, h);
        const double dAd = dotb(nb, begini, off, d, Ad);
        const double alpha = (fabs(rth) < 1.e-20 || fabs(rth) > 1.e20) ? 0. : rth / dAd;
        double m = 0.;
        for (int i = begini; i < nb; ++i) {
            for (int j = off[i]; j < off[i + 1]; ++j) {
                x[j] += alpha * d[j];
                const double dr = alpha * Ad[j];
                r1[j] = r[j] - dr;
                m += dr * dr;
            }
        }
        precond(nb, begini, off, Sb, Sf, Cinv, omega, r1, h1, tw, uw);
        if (k == 0) m0 = m;
        if (trace_this && k < g_pcg_trace_cap) {
            g_pcg_trace[k] = m / m0;
            g_pcg_trace_len = k + 1;
        }
        if (m < mold) {
            memcpy(xo, x, sizeof(double) * (size_t)n);
            mold = m;
        }
        if (m * 1.e5 < m0) break;
        for (int i = begini; i < nb; ++i)
            for (int j = off[i]; j < off[i + 1]; ++j) hd[j] = h1[j] - h[j];
        const double num = dotb(nb, begini, off, r1, hd);
        const double den = dotb(nb, begini, off, r, h);
        double beta = (fabs(num) < 1.e-10) ? 0. : num / den;
        if (beta < 0.) beta = 0.;
        for (int i = begini; i < nb; ++i)
            for (int j = off[i]; j < off[i + 1]; ++j) {
                d[j] = h1[j] + beta * d[j];
                r[j] = r1[j];
                h[j] = h1[j];
            }
    }
    /* entries of blocks below begini keep the copy of b (exivold = exiv) */
    memcpy(xout, xo, sizeof(double) * (size_t)n);
    free(x); free(xo); free(r); free(r1); free(d); free(Ad); free(h); free(h1); free(hd); free(Cinv);
    free(tw); free(uw);
    return (k < msize) ? k + 1 : k;
}

int oracle_block_pcg(int nb, int begini, const int *off, double *const *Sb, const unsigned char *Sf,
                     const double *b, double *xout)
{
    return oracle_block_pcg_w(nb, begini, off, Sb, Sf, b, xout, 0.);
}

/* Dense Cholesky solve of a symmetric positive definite system (row-major
 * A[n*n], overwritten by L in its lower triangle): A = L L^T by the
 * Cholesky-Crout recurrence in column order, then L y = b, L^T x = y.  This is
 * the exact solve of the reduced camera system used by the dense form
 * (solver = 1); the reference solves the same system with SparsePCG
 * (rbacore.cpp:1807-1914).  Returns 0, or -1 at the first pivot <= 0. */
int oracle_chol_solve(int n, double *A, const double *b, double *x)
{
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.)) return -1;
        const double ljj = sqrt(d);
        A[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / ljj;
        }
    }
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= A[(size_t)i * n + k] * x[k];
        x[i] = s / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s -= A[(size_t)k * n + i] * x[k];
        x[i] = s / A[(size_t)i * n + i];
    }
    return 0;
}

/* The dense form's solve of S x = b over the touched blocks (solver = 1):
 * S assembled densely from the upper blocks, both triangles, then
 * oracle_chol_solve.  A non-positive pivot gives a NaN step, which the
 * caller's NaN check turns into the reference's error path. */
static void dense_solve(orc *o, double *sol)
{
    const int nc = o->nimg, n = o->nrioff;
    double *A = (double *)xcalloc((size_t)n * n + 1, sizeof(double));
    for (int k = 0; k < nc; ++k)
        for (int l = k; l < nc; ++l) {
            const size_t id = (size_t)k * nc + l;
            if (!o->Sf[id]) continue;
            for (int r = 0; r < bsz(o, k); ++r)
                for (int c = 0; c < bsz(o, l); ++c) {
                    const double v = o->Sb[id][r * MAXB + c];
                    A[(size_t)(o->camparv[k] + r) * n + o->camparv[l] + c] = v;
                    A[(size_t)(o->camparv[l] + c) * n + o->camparv[k] + r] = v;
                }
        }
    if (oracle_chol_solve(n, A, o->ex, sol))
        for (int j = 0; j < n; ++j) sol[j] = NAN;
    free(A);
}

/* CalcBundleSep (rbacore.cpp:1094-1195) */
static int bundle_sep(orc *o, double mult)
{
    const double mult1 = 1. + mult;
    const int nc = o->nimg;
    o->begini = (o->aoflag || o->nraddpi > 0) ? 0 : 1; /* rbacore.cpp:1104-1106 */
    for (size_t id = 0; id < (size_t)nc * nc; ++id) {
        o->Sf[id] = 0;
        if (o->Sb[id]) memset(o->Sb[id], 0, sizeof(double) * MAXB * MAXB);
    }
    memset(o->ex, 0, sizeof(double) * (size_t)o->nrioff);
    for (int i = o->begini; i < nc; ++i) sblock(o, i, i); /* diagonal always present */
    corrections(o, 1, mult1, o->XP, 0);
    for (int i = o->begini; i < nc; ++i) {
        double *D = o->Sb[(size_t)i * nc + i];
        const int m = bsz(o, i);
        for (int j = 0; j < m; ++j) D[j * MAXB + j] *= mult1;
        for (int j = 0; j < m; ++j) D[j * MAXB + j] += 1.e-2;
    }
    double *sol = (double *)xcalloc((size_t)o->nrioff + 1, sizeof(double));
    if (o->solver == 1) {
        dense_solve(o, sol);
        o->pcg_last = 0;
    } else {
        o->pcg_last = oracle_block_pcg_w(nc, o->begini, o->camparv, o->Sb, o->Sf, o->ex, sol, o->omega);
    }
    o->pcg_total += o->pcg_last;
    memcpy(o->ex, sol, sizeof(double) * (size_t)o->nrioff);
    free(sol);
    for (int j = o->begini; j < nc; ++j)
        for (int k = o->camparv[j]; k < o->camparv[j + 1]; ++k) o->x[k] = (float)o->ex[k]; /* H3 */
    corrections(o, 2, mult1, o->XP, 0);
    return 0;
}

/* Reweight2D (rbacore.cpp:823-887) */
static void reweight(orc *o, int iflag)
{
    const int nc = o->nimg;
    float *key = (float *)xcalloc((size_t)o->ntr + 1, sizeof(float));
    /* keys in camera-major order: image k occupies [nrobvo[k], nrobvo[k+1]) */
    for (int i = 0; i < o->npts; ++i) {
        int pxs = (o->toff[i + 1] - o->toff[i]) * 2;
        if (pxs < 4) pxs = 4;
        const float rf = (float)pxs / (float)(pxs - 3);
        const float redfac = rf * rf;
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q) key[o->nro[q]] = o->sig0rmv[o->nro[q]] * redfac;
    }
    float *s00 = (float *)xcalloc((size_t)nc, sizeof(float));
    for (int k = 0; k < nc; ++k) {
        const int b = o->nrobvo[k], n = o->nrobvo[k + 1] - b;
        s00[k] = 1.f / (oracle_upper_median(key + b, n) * 2.1981f);
        if (iflag > 1) s00[k] *= 0.001f;
    }
    double ws = 0.;
    for (int q = 0; q < o->ntr; ++q) {
        const int k = o->timg[q], nro = o->nro[q];
        float ww = o->sig0rmv[nro] * s00[k];
        o->ISflag[nro] = ww > 9.f; /* kf, rpba.h:529 */
        if (o->ISflag[nro]) {
            ww = 1.e-4f;
            o->nrobs -= 2;
        } else
            ww = 1.f;
        float *W = o->Wij + 3 * nro;
        const float *w0 = o->wold + 3 * q;
        W[0] = ww * w0[0]; W[1] = ww * w0[1]; W[2] = ww * w0[2];
        if (W[0] + W[1] < 2.e-8f) { W[0] = W[1] = 1.e-8f; W[2] = 0.f; }
        ws = o->fsites ? (double)((float)ws + (W[0] + W[1])) : ws + ((double)W[0] + (double)W[1]);
    }
    o->wsum = ws;
    free(key);
    free(s00);
}

/* StoreResultsIteration (rbacore.cpp:891-918) */
static void store(orc *o)
{
    memcpy(o->XP, o->XPn, sizeof(double) * o->xpsize);
    memcpy(o->Rs, o->Rn, sizeof(double) * 9 * (size_t)o->nimg);
    memcpy(o->Cs, o->Cn, sizeof(double) * 3 * (size_t)o->nimg);
    for (int k = 0; k < o->nimg; ++k) {
        double *P = o->PM + 12 * k;
        const double *R = o->Rn + 9 * k, *C = o->Cn + 3 * k;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) P[r * 4 + c] = R[r * 3 + c];
            P[r * 4 + 3] = R[r * 3 + 0] * C[0] + R[r * 3 + 1] * C[1] + R[r * 3 + 2] * C[2];
        }
        o->cf[k] = o->nf[k]; o->cd0[k] = o->nd0[k]; o->cd1[k] = o->nd1[k];
    }
    memcpy(o->XX, o->XXn, sizeof(double) * 3 * (size_t)o->npts);
    memcpy(o->ISflagnew, o->ISflag, (size_t)o->ntr);
    o->nrobsnew = o->nrobs;
    o->nrunknownsnew = o->nrunknowns;
    o->redundancynew = o->redundancy;
}

/* RobustDelete (rbacore.cpp:922-995): drop flagged observations; drop a point
 * whose remaining track is shorter than minimagenrdel */
static void robust_delete(orc *o)
{
    int *ntoff = (int *)xcalloc((size_t)o->npts + 1, sizeof(int));
    int np2 = 0, nt2 = 0;
    for (int i = 0; i < o->npts; ++i) {
        int nf = 0;
        const int t = o->toff[i + 1] - o->toff[i];
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q) nf += o->ISflagnew[o->nro[q]] ? 1 : 0;
        const int left = t - nf;
        if (nf > 0 && (g_mut == 4 ? left <= o->minimagenrdel : left < o->minimagenrdel)) continue; /* point erased */
        const int b = nt2;
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q) {
            if (nf > 0 && o->ISflagnew[o->nro[q]]) continue;
            o->timg[nt2] = o->timg[q];
            o->tfo[nt2] = o->tfo[q];
            o->wold[3 * nt2 + 0] = o->wold[3 * q + 0];
            o->wold[3 * nt2 + 1] = o->wold[3 * q + 1];
            o->wold[3 * nt2 + 2] = o->wold[3 * q + 2];
            ++nt2;
        }
        ntoff[np2] = b;
        o->nrpv[np2] = o->nrpv[i];
        for (int j = 0; j < 3; ++j) o->XX[3 * np2 + j] = o->XX[3 * i + j];
        ++np2;
    }
    ntoff[np2] = nt2;
    memcpy(o->toff, ntoff, sizeof(int) * (size_t)(np2 + 1));
    free(ntoff);
    o->npts = np2;
    o->ntr = nt2;
}

/* ------------------------------------------------------------------------- */
/* Adjust (rbacore.cpp:333-616), non-GCP, itbreakflag = false                */
/* ------------------------------------------------------------------------- */

static double fsq(const orc *o, double v) { return o->fsites ? (double)sqrtf((float)v) : sqrt(v); }

/* position-weighted sum: equal states give equal values, different ones
 * (practically always) different values */
static double fp_add(double acc, const double *v, size_t n, size_t base)
{
    for (size_t i = 0; i < n; ++i) acc += v[i] * (1. + 1.e-3 * (double)((base + i) % 977));
    return acc;
}
static double fp_cams(const orc *o, const double *R, const double *Cc, const double *f, const double *d0,
                      const double *d1)
{
    const size_t n = (size_t)o->nimg;
    double a = fp_add(0., R, 9 * n, 0);
    a = fp_add(a, Cc, 3 * n, 9 * n);
    a = fp_add(a, f, n, 12 * n);
    a = fp_add(a, d0, n, 13 * n);
    return fp_add(a, d1, n, 14 * n);
}
static double fp_keys(const orc *o)
{
    double a = 0.;
    for (int q = 0; q < o->ntr; ++q) a += (double)o->sig0rmv[q] * (1. + 1.e-3 * (double)(q % 977));
    return a;
}
static double fp_flags(const unsigned char *f, int n)
{
    double c = 0.;
    for (int q = 0; q < n; ++q) c += f[q] ? (double)(q + 1) : 0.;
    return c;
}
/* the next scripted sig0, if a script is set */
static void script_sig0(orc *o)
{
    if (o->script && o->script_pos < o->script_len) o->sig0 = o->script[o->script_pos++];
}

void orc_set_script(orc *o, int n, const double *seq)
{
    o->script = seq;
    o->script_len = n;
    o->script_pos = 0;
}
void orc_set_rec(orc *o, orc_rec *rec) { o->rec = rec; }

/* RobustDelete alone on the current tracks, with the labels ISflagnew given
 * per track entry (point-major order q).  Writes the original indices of the
 * kept points (kept_pts[0..np')) and the file-order indices of the kept
 * observations (kept_obs[0..nt')), and np', nt' into counts. */
void orc_delete_with_flags(orc *o, const unsigned char *flags_q, int *kept_pts, int *kept_obs, int *counts)
{
    for (int q = 0; q < o->ntr; ++q) o->ISflagnew[o->nro[q]] = flags_q[q] ? 1 : 0;
    robust_delete(o);
    for (int i = 0; i < o->npts; ++i) kept_pts[i] = o->nrpv[i];
    for (int q = 0; q < o->ntr; ++q) kept_obs[q] = o->tfo[q];
    counts[0] = o->npts;
    counts[1] = o->ntr;
}
/* intrinsics (f, d0, d1) per image: which = 0 stored (cameras), 1 last trial
 * (camerasnew) */
void orc_cams(const orc *o, int which, double *out)
{
    for (int k = 0; k < o->nimg; ++k) {
        out[3 * k] = which ? o->nf[k] : o->cf[k];
        out[3 * k + 1] = which ? o->nd0[k] : o->cd0[k];
        out[3 * k + 2] = which ? o->nd1[k] : o->cd1[k];
    }
}

int orc_adjust(orc *o, rpba_log_fn log, void *user, rpba_stats *st)
{
    const int robust = o->robust;
    int itt = 0, ittflag = 1;
    const double frac = o->lflflag ? (double)1.01f : (double)1.05f;
    double mult = 1.e-3;
    int rc = 0;
    char buf[512];
    if (st) { st->passes_done = 0; st->trace_len = 0; }
    for (int it = 0; it < 2; ++it) {
        o->sig0old = 1.e20f;
        if (it > 0) mult = 1.e-6;
        o->imagepointnr = o->ntr;
        o->nrobs = 2 * o->imagepointnr;
        rc = count_images(o, log, user);
        if (rc) break;
        if (ittflag) {
            memset(o->ISflag, 0, (size_t)o->ntr);
            memset(o->ISflagnew, 0, (size_t)o->ntr);
            prepare_weights(o);
            ittflag = 0;
        }
        o->wsum = o->fsites ? (double)((float)(2 * o->imagepointnr) / (float)o->wsum) : (double)(2 * o->imagepointnr) / o->wsum;
        o->wsumg = o->wsum;
        for (int i = 0; i < o->npts; ++i)
            for (int j = 0; j < 3; ++j) o->XP[o->nrioff + 3 * i + j] = o->XX[3 * i + j];
        encode_cameras(o);
        o->nrunknowns = o->nrioff3 + o->nraddpi * o->nimg; /* H7 */
        memcpy(o->XXn, o->XX, sizeof(double) * 3 * (size_t)o->npts);
        memset(o->x, 0, sizeof(double) * o->xpsize);
        o->redundancy = o->nrobs - o->nrunknowns;
        if (o->redundancy < 0) {
            if (o->nimg > 3) orc_log(log, user, "BA-Error: Low redundancy\n");
            rc = RPBA_E_LOW_REDUNDANCY;
            break;
        }
        if (it == 0) {
            decode_cameras(o, o->XP);
            corrections(o, 0, 1. + mult, o->XP, 1);
            o->sig0 = FS(o, FS(o, o->sig0 * o->wsumg) / (double)(float)o->redundancy);
            script_sig0(o);
            o->sig0in = o->sig0old = o->sig0;
            memcpy(o->Rs, o->Rn, sizeof(double) * 9 * (size_t)o->nimg);
            memcpy(o->Cs, o->Cn, sizeof(double) * 3 * (size_t)o->nimg);
            if (o->rec) o->rec->init_key = fp_keys(o);
        } else {
            o->sig0in = o->sig0old;
        }
        int iflag = robust ? 2 : 1;
        int iter = 1;
        const int nrunknownsgiven = o->nrunknowns, nrobsgiven = o->nrobs;
        while (iflag > 0) {
            orc_rec *rc_ = (o->rec && o->rec->len < o->rec->cap) ? o->rec : NULL;
            const int ri = rc_ ? rc_->len : 0;
            if (rc_) rc_->key_used[ri] = (it == 0 && robust) ? fp_keys(o) : 0.;
            if (it == 0 && robust) {
                o->nrunknowns = nrunknownsgiven;
                o->nrobs = nrobsgiven;
                reweight(o, iflag);
                o->redundancy = o->nrobs - o->nrunknowns;
                o->wsum = o->fsites ? (double)((float)(2 * o->imagepointnr) / (float)o->wsum) : (double)(2 * o->imagepointnr) / o->wsum;
                o->wsumg = o->wsum;
            }
            if (rc_) {
                rc_->lin_cam[ri] = fp_cams(o, o->Rn, o->Cn, o->nf, o->nd0, o->nd1);
                rc_->lin_pts[ri] = fp_add(0., o->XP + o->nrioff, 3 * (size_t)o->npts, 0);
            }
            bundle_sep(o, mult);
            for (int i = 0; i < o->nrunknowns; ++i) {
                if (isnan(o->x[i]) || isinf(o->x[i])) {
                    snprintf(buf, sizeof buf, "BA-Error: NANs in result vector %d %g %d %d\n", i, o->x[i], o->nrioff, o->nrioff3);
                    orc_log(log, user, buf);
                    rc = RPBA_E_NAN_RESULT;
                    break;
                }
                o->XPn[i] = o->XP[i] - o->x[i];
            }
            if (rc) break;
            decode_cameras(o, o->XPn);
            corrections(o, 0, 1. + mult, o->XPn, 1);
            o->sig0 = FS(o, FS(o, o->sig0 * o->wsumg) / (double)(float)o->redundancy);
            script_sig0(o);
            if (rc_) {
                rc_->trial_cam[ri] = fp_cams(o, o->Rn, o->Cn, o->nf, o->nd0, o->nd1);
                rc_->trial_pts[ri] = fp_add(0., o->XPn + o->nrioff, 3 * (size_t)o->npts, 0);
                rc_->trial_key[ri] = fp_keys(o);
            }
            int acc = 0;
            if (o->sig0old < FS(o, o->sig0 * frac)) --iflag;
            if (o->sig0 <= o->sig0old || (iter == 1 && g_mut != 2)) {
                store(o);
                o->sig0old = o->sig0;
                acc = 1;
            } else if (g_mut == 1) {
                memcpy(o->Rn, o->Rs, sizeof(double) * 9 * (size_t)o->nimg);
                memcpy(o->Cn, o->Cs, sizeof(double) * 3 * (size_t)o->nimg);
                memcpy(o->nf, o->cf, sizeof(double) * (size_t)o->nimg);
                memcpy(o->nd0, o->cd0, sizeof(double) * (size_t)o->nimg);
                memcpy(o->nd1, o->cd1, sizeof(double) * (size_t)o->nimg);
            }
            if (!acc && g_mut == 3) memcpy(o->ISflagnew, o->ISflag, (size_t)o->ntr);
            if (rc_) {
                rc_->stored_cam[ri] = fp_cams(o, o->Rs, o->Cs, o->cf, o->cd0, o->cd1);
                rc_->stored_pts[ri] = fp_add(0., o->XX, 3 * (size_t)o->npts, 0);
                rc_->flag_fp[ri] = fp_flags(o->ISflag, o->ntr);
                rc_->flagnew_fp[ri] = fp_flags(o->ISflagnew, o->ntr);
                rc_->iflag[ri] = iflag;
                rc_->pass[ri] = it;
                rc_->len++;
            }
            if (st && st->trace_sig0 && st->trace_len < st->trace_cap) {
                st->trace_sig0[st->trace_len] = o->sig0;
                if (st->trace_accept) st->trace_accept[st->trace_len] = acc;
                if (st->trace_pcg) st->trace_pcg[st->trace_len] = o->pcg_last;
                if (st->trace_nrobs) st->trace_nrobs[st->trace_len] = o->nrobs;
                st->trace_len++;
            }
            if (g_mut == 5 ? itt >= 200 : itt > 200) break;
            ++itt;
            if (g_stop_after > 0 && itt >= g_stop_after) break;
            ++iter;
        }
        if (rc) break;
        const double s_in = fsq(o, FS(o, o->sig0in * (double)o->ihw2));
        const double s_out = fsq(o, FS(o, o->sig0old * (double)o->ihw2));
        int red, nob, nun;
        if (it == 0 && robust) { red = o->redundancynew; nob = o->nrobsnew; nun = o->nrunknownsnew; }
        else { red = o->redundancy; nob = o->nrobs; nun = o->nrunknowns; }
        snprintf(buf, sizeof buf, "Sigma0: %g %g  %d %d %d  IT: %d  ITT: %d\n", s_in, s_out, red, nob, nun, it, itt);
        orc_log(log, user, buf);
        if (st) {
            st->sigma_in[it] = s_in;
            st->sigma_out[it] = s_out;
            st->redundancy[it] = red;
            st->nobs[it] = nob;
            st->nunknowns[it] = nun;
            st->passes_done = it + 1;
        }
        if (robust && it == 0) robust_delete(o);
        ittflag = 1;
        if (g_stop_after > 0 && itt >= g_stop_after) break;
    }
    if (st) {
        st->itt = itt;
        st->pcg_iters = o->pcg_total;
        st->sig0_final = o->sig0old;
    }
    o->lasterr = rc;
    return rc;
}

/* OutputLSAc + PointXXvOut (rbacore.cpp:644-666, 1978-2010).  Camera
 * intrinsics come from the last trial (camerasnew), R|t and X from the stored
 * state. */
void orc_output(orc *o, rpba_block *io)
{
    if (io->obs_kept) memset(io->obs_kept, 0, (size_t)io->n_obs);
    if (io->pt_kept) memset(io->pt_kept, 0, (size_t)io->n_pts);
    for (int i = 0; i < o->npts; ++i) {
        const int p = o->nrpv[i];
        for (int j = 0; j < 3; ++j) io->pts[3 * p + j] = (float)o->XX[3 * i + j];
        if (io->pt_kept) io->pt_kept[p] = 1;
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q)
            if (io->obs_kept) io->obs_kept[o->tfo[q]] = 1;
    }
    for (int k = 0; k < o->nimg; ++k) {
        for (int j = 0; j < 12; ++j) io->cam_P[12 * k + j] = (float)o->PM[12 * k + j];
        io->cam_f[k] = (float)(g_mut == 6 ? o->cf[k] : o->nf[k]);
        io->cam_d[2 * k] = (float)(g_mut == 6 ? o->cd0[k] : o->nd0[k]);
        io->cam_d[2 * k + 1] = (float)(g_mut == 6 ? o->cd1[k] : o->nd1[k]);
    }
}

/* EvalLSA + CalcCorrPointMulti (rpbacore.cpp:882-941, 778-837) on the given
 * parameters; returns sqrt(sig0g*ihw2) ("ITTTT:" line). */
double oracle_exteval(const rpba_block *io, const rpba_flags *f, int fsites, int64_t *nobs_out, int64_t *dof_out)
{
    const int nc = io->n_cams, np = io->n_pts, no = io->n_obs;
    const int nraddp = (f->est_k2 != 0) + (f->est_k4 != 0) + (f->est_f != 0);
    float ihw2 = 0.f;
    for (int i = 0; i < nc; ++i) {
        const int w = io->img_w ? io->img_w[i] : 1, h = io->img_h ? io->img_h[i] : 1;
        ihw2 += (float)((h + w) * (h + w)) * 0.25f;
    }
    ihw2 /= (float)nc;
    int *cnt = (int *)xcalloc((size_t)np + 1, sizeof(int));
    for (int i = 0; i < no; ++i)
        if (io->obs_cam[i] >= 0) cnt[io->obs_pt[i] + 1]++;
    for (int i = 0; i < np; ++i) cnt[i + 1] += cnt[i];
    int *all = (int *)xcalloc((size_t)cnt[np] + 1, sizeof(int));
    int *pos = (int *)xcalloc((size_t)np + 1, sizeof(int));
    memcpy(pos, cnt, sizeof(int) * (size_t)(np + 1));
    for (int i = 0; i < no; ++i)
        if (io->obs_cam[i] >= 0) all[pos[io->obs_pt[i]]++] = i;
    long red = 0, nobs = 0;
    double sg = 0.;
    for (int i = 0; i < np; ++i) {
        const int t = cnt[i + 1] - cnt[i];
        if (t <= 1) continue;
        double ws = 0.;
        for (int j = cnt[i]; j < cnt[i + 1]; ++j) {
            const int q = all[j];
            const float wx = io->obs_w ? io->obs_w[3 * q] : 1.f, wy = io->obs_w ? io->obs_w[3 * q + 1] : 1.f;
            ws = fsites ? (double)((float)ws + (wx + wy)) : ws + ((double)wx + (double)wy);
        }
        ws = fsites ? (double)((float)ws / (float)(2 * t)) : ws / (double)(2 * t);
        const double X0 = io->pts[3 * i], X1 = io->pts[3 * i + 1], X2 = io->pts[3 * i + 2];
        double s = 0.;
        for (int j = cnt[i]; j < cnt[i + 1]; ++j) {
            const int q = all[j], c = io->obs_cam[q];
            const float *P = io->cam_P + 12 * c;
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
            const double z = 1. + ((double)io->cam_d[2 * c] * dist + (double)io->cam_d[2 * c + 1] * dist * dist);
            p1 *= z;
            p2 *= z;
            p1 *= (double)io->cam_f[c];
            p2 *= (double)io->cam_f[c];
            const double dx = (double)io->obs_xy[2 * q] - p1, dy = (double)io->obs_xy[2 * q + 1] - p2;
            const float wx = io->obs_w ? io->obs_w[3 * q] : 1.f, wy = io->obs_w ? io->obs_w[3 * q + 1] : 1.f;
            const float wxy = io->obs_w ? io->obs_w[3 * q + 2] : 0.f;
            const double e = dx * dx * wx + dy * dy * wy + 2 * dx * dy * wxy;
            s = fsites ? (double)((float)s + (float)e) : s + e;
        }
        sg = fsites ? (double)((float)sg + (float)s / (float)ws) : sg + s / ws;
        red += 2 * t - 3;
        nobs += 2 * t;
    }
    const long dof = red - (long)(nc - 2) * 6 - 5 - (long)nc * nraddp;
    sg = fsites ? (double)((float)sg * (1.f / (float)dof)) : sg * (1. / (double)dof);
    free(cnt); free(all); free(pos);
    if (nobs_out) *nobs_out = nobs;
    if (dof_out) *dof_out = dof;
    return fsites ? (double)sqrtf((float)sg * ihw2) : sqrt(sg * (double)ihw2);
}

/* CalcPLSA restated for the paths this oracle covers (rpbacore.cpp:28-99,
 * 511-527). */
int oracle_calc_plsa(rpba_block *io, const rpba_flags *f, int fsites, rpba_log_fn log, void *user, rpba_stats *st)
{
    if (f->exteval) {
        int64_t nobs = 0, dof = 0;
        const double s = oracle_exteval(io, f, fsites, &nobs, &dof);
        char buf[160];
        snprintf(buf, sizeof buf, "ITTTT: %g %lld %lld\n", s, (long long)nobs, (long long)dof);
        orc_log(log, user, buf);
        if (st) { st->exteval_sigma = s; st->exteval_nobs = nobs; st->exteval_dof = dof; }
        return 0;
    }
    orc *o = orc_create(io, f, fsites);
    const int rc = orc_adjust(o, log, user, st);
    orc_output(o, io);
    if (st) {
        st->n_pts_kept = o->npts;
        st->n_obs_kept = o->ntr;
    }
    orc_destroy(o);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* test hooks                                                                */
/* ------------------------------------------------------------------------- */

/* Pass-0 setup of Adjust up to the initial cost (rbacore.cpp:370-499) so the
 * tests can probe one linearisation. */
int orc_setup_pass0(orc *o)
{
    o->imagepointnr = o->ntr;
    o->nrobs = 2 * o->ntr;
    int rc = count_images(o, NULL, NULL);
    if (rc) return rc;
    memset(o->ISflag, 0, (size_t)o->ntr);
    memset(o->ISflagnew, 0, (size_t)o->ntr);
    prepare_weights(o);
    o->wsum = (double)(2 * o->imagepointnr) / o->wsum;
    o->wsumg = o->wsum;
    for (int i = 0; i < o->npts; ++i)
        for (int j = 0; j < 3; ++j) o->XP[o->nrioff + 3 * i + j] = o->XX[3 * i + j];
    encode_cameras(o);
    o->nrunknowns = o->nrioff3 + o->nraddpi * o->nimg;
    memcpy(o->XXn, o->XX, sizeof(double) * 3 * (size_t)o->npts);
    decode_cameras(o, o->XP);
    corrections(o, 0, 1.001, o->XP, 1);
    return 0;
}

/* CPU-baseline sample of one LM iteration of pass 0 (timing only; its
 * numbers are not a solve): the per-point work of CalcCorrections -- the
 * linearisation with the Schur pair loop, the back-substitution and the trial
 * cost (rbacore.cpp:1201-1410) -- over every stride-th point, then reps S d
 * products (eAAdv, rbacore.cpp:2048-2068) over the blocks that sample formed.
 * out: [0] s linearise + Schur, [1] s back-substitution, [2] s cost, [3] s per
 * stored block per S d product, [4] stored blocks (diagonal included), [5]
 * points sampled, [6] observations sampled, [7] point pairs sampled. */
int orc_sample_iteration(orc *o, int stride, int reps, double *out)
{
    struct timespec t0, t1;
#define ORC_NOW(t) clock_gettime(CLOCK_MONOTONIC, &(t))
#define ORC_SEC(a, b) ((double)((b).tv_sec - (a).tv_sec) + 1e-9 * (double)((b).tv_nsec - (a).tv_nsec))
    o->imagepointnr = o->ntr;
    o->nrobs = 2 * o->ntr;
    int rc = count_images(o, NULL, NULL);
    if (rc) return rc;
    prepare_weights(o);
    for (int i = 0; i < o->npts; ++i)
        for (int j = 0; j < 3; ++j) o->XP[o->nrioff + 3 * i + j] = o->XX[3 * i + j];
    encode_cameras(o);
    o->nrunknowns = o->nrioff3 + o->nraddpi * o->nimg;
    decode_cameras(o, o->XP);
    o->pt_stride = stride;
    const int nc = o->nimg;
    o->begini = (o->aoflag || o->nraddpi > 0) ? 0 : 1; /* rbacore.cpp:1104-1106 */
    for (int i = o->begini; i < nc; ++i) sblock(o, i, i);
    ORC_NOW(t0);
    corrections(o, 1, 1.001, o->XP, 0);
    ORC_NOW(t1);
    out[0] = ORC_SEC(t0, t1);
    ORC_NOW(t0);
    corrections(o, 2, 1.001, o->XP, 0);
    ORC_NOW(t1);
    out[1] = ORC_SEC(t0, t1);
    ORC_NOW(t0);
    corrections(o, 0, 1.001, o->XP, 1);
    ORC_NOW(t1);
    out[2] = ORC_SEC(t0, t1);
    long long nblk = 0;
    for (size_t id = 0; id < (size_t)nc * nc; ++id) nblk += o->Sf[id] ? 1 : 0;
    double *d = (double *)xcalloc((size_t)o->nrioff + 1, sizeof(double));
    double *y = (double *)xcalloc((size_t)o->nrioff + 1, sizeof(double));
    for (int j = 0; j < o->nrioff; ++j) d[j] = 1.e-3 * (double)(j % 17);
    ORC_NOW(t0);
    for (int r = 0; r < reps; ++r) spmv(nc, o->begini, o->camparv, o->Sb, o->Sf, d, y);
    ORC_NOW(t1);
    out[3] = ORC_SEC(t0, t1) / ((double)(reps > 0 ? reps : 1) * (double)(nblk > 0 ? nblk : 1));
    out[4] = (double)nblk;
    long long np = 0, no = 0, npair = 0;
    for (int i = 0; i < o->npts; i += (stride > 1 ? stride : 1)) {
        const long long t = o->toff[i + 1] - o->toff[i];
        ++np;
        no += t;
        npair += t * (t - 1) / 2;
    }
    out[5] = (double)np;
    out[6] = (double)no;
    out[7] = (double)npair;
    free(d);
    free(y);
    o->pt_stride = 0;
    return 0;
#undef ORC_NOW
#undef ORC_SEC
}

/* dims: nimg, npts, ntr, nrioff, nrioff3, normcamera, esvmaxm, nraddpi */
void orc_dims(const orc *o, int *out)
{
    out[0] = o->nimg; out[1] = o->npts; out[2] = o->ntr; out[3] = o->nrioff;
    out[4] = o->nrioff3; out[5] = o->normcamera; out[6] = o->esvmaxm; out[7] = o->nraddpi;
}

void orc_camparv(const orc *o, int *out)
{
    for (int i = 0; i <= o->nimg; ++i) out[i] = o->camparv[i];
}

void orc_get_theta(const orc *o, double *theta) { memcpy(theta, o->XP, sizeof(double) * (size_t)o->nrioff3); }

/* residuals (point-major track order) and float keys at parameters theta;
 * the trial cameras are left at theta (like the reference's DecodeCameras) */
void orc_residuals(orc *o, const double *theta, double *r, float *key)
{
    decode_cameras(o, theta);
    corrections(o, 0, 1.001, theta, 1);
    for (int i = 0; i < o->npts; ++i) {
        const double *X = theta + o->nrioff + 3 * i;
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q) {
            obs_eval_t e;
            eval_obs(o, o->timg[q], X, o->obs_xy + 2 * o->tfo[q], &e);
            r[2 * q] = e.diff1;
            r[2 * q + 1] = e.diff2;
            if (key) key[q] = o->sig0rmv[o->nro[q]];
        }
    }
    o->sig0 = o->sig0;
}

/* per-track-entry Jacobians at the current trial cameras and points XXn:
 * Bc [ntr][2][MAXB] (compact camera columns), Bp [ntr][2][3] */
void orc_jacobians(orc *o, double *Bc, double *Bp, int *ncols)
{
    for (int i = 0; i < o->npts; ++i)
        for (int q = o->toff[i]; q < o->toff[i + 1]; ++q) {
            obs_eval_t e;
            eval_obs(o, o->timg[q], o->XXn + 3 * i, o->obs_xy + 2 * o->tfo[q], &e);
            double bc[2 * MAXB] = {0}, bp[6];
            ncols[q] = derivs(o, o->timg[q], &e, bc, bp);
            memcpy(Bc + (size_t)q * 2 * MAXB, bc, sizeof bc);
            memcpy(Bp + (size_t)q * 6, bp, sizeof bp);
        }
}

/* One linearisation + Schur at the current state (first half of CalcBundleSep
 * with the S damping): dense S [nrioff^2], b [nrioff]. */
void orc_linearize(orc *o, double mult, double *Sd, double *b)
{
    const int nc = o->nimg, n = o->nrioff;
    const double mult1 = 1. + mult;
    o->begini = (o->aoflag || o->nraddpi > 0) ? 0 : 1; /* rbacore.cpp:1104-1106 */
    for (size_t id = 0; id < (size_t)nc * nc; ++id) {
        o->Sf[id] = 0;
        if (o->Sb[id]) memset(o->Sb[id], 0, sizeof(double) * MAXB * MAXB);
    }
    memset(o->ex, 0, sizeof(double) * (size_t)n);
    for (int i = o->begini; i < nc; ++i) sblock(o, i, i);
    corrections(o, 1, mult1, o->XP, 0);
    for (int i = o->begini; i < nc; ++i) {
        double *D = o->Sb[(size_t)i * nc + i];
        const int m = bsz(o, i);
        for (int j = 0; j < m; ++j) D[j * MAXB + j] *= mult1;
        for (int j = 0; j < m; ++j) D[j * MAXB + j] += 1.e-2;
    }
    memset(Sd, 0, sizeof(double) * (size_t)n * n);
    for (int k = 0; k < nc; ++k)
        for (int l = k; l < nc; ++l) {
            const size_t id = (size_t)k * nc + l;
            if (!o->Sf[id]) continue;
            for (int r = 0; r < bsz(o, k); ++r)
                for (int c = 0; c < bsz(o, l); ++c) {
                    const double v = o->Sb[id][r * MAXB + c];
                    Sd[(size_t)(o->camparv[k] + r) * n + o->camparv[l] + c] = v;
                    Sd[(size_t)(o->camparv[l] + c) * n + o->camparv[k] + r] = v;
                }
        }
    memcpy(b, o->ex, sizeof(double) * (size_t)n);
}

/* One full CalcBundleSep at the current state: x [nrioff3]; returns PCG its */
int orc_step(orc *o, double mult, double *x)
{
    bundle_sep(o, mult);
    memcpy(x, o->x, sizeof(double) * (size_t)o->nrioff3);
    return o->pcg_last;
}

/* PCG on a dense symmetric matrix cut into blocks of the given sizes */
int oracle_pcg_dense_w(int nb, const int *bs, const double *Sd, const double *b, double *x, double omega);
int oracle_pcg_dense(int nb, const int *bs, const double *Sd, const double *b, double *x)
{
    return oracle_pcg_dense_w(nb, bs, Sd, b, x, 0.);
}

int oracle_pcg_dense_w(int nb, const int *bs, const double *Sd, const double *b, double *x, double omega)
{
    int *off = (int *)xcalloc((size_t)nb + 1, sizeof(int));
    for (int i = 0; i < nb; ++i) off[i + 1] = off[i] + bs[i];
    const int n = off[nb];
    double **Sb = (double **)xcalloc((size_t)nb * nb, sizeof(double *));
    unsigned char *Sf = (unsigned char *)xcalloc((size_t)nb * nb, 1);
    for (int k = 0; k < nb; ++k)
        for (int l = k; l < nb; ++l) {
            int nz = (k == l);
            for (int r = 0; r < bs[k]; ++r)
                for (int c = 0; c < bs[l]; ++c)
                    if (Sd[(size_t)(off[k] + r) * n + off[l] + c] != 0.) nz = 1;
            if (!nz) continue;
            double *B = (double *)xcalloc(MAXB * MAXB, sizeof(double));
            for (int r = 0; r < bs[k]; ++r)
                for (int c = 0; c < bs[l]; ++c) B[r * MAXB + c] = Sd[(size_t)(off[k] + r) * n + off[l] + c];
            Sb[(size_t)k * nb + l] = B;
            Sf[(size_t)k * nb + l] = 1;
        }
    const int it = oracle_block_pcg_w(nb, 0, off, Sb, Sf, b, x, omega);
    for (size_t i = 0; i < (size_t)nb * nb; ++i) free(Sb[i]);
    free(Sb); free(Sf); free(off);
    return it;
}

/* the partitioned solve (rpbacore.cpp:100-510, 536-877) */
#include "rpba_oracle_part.c"
