/*
 * rpba_oracle_part.c -- the partitioned RPBA solve of CalcPLSA
 * (rpbacore.cpp:100-510): per-part LSA blocks in GCP mode (one damped
 * Gauss-Newton step each, rbacore.cpp:233-329, 333-616 with itbreakflag,
 * 620-640, 1677-1729) alternating with the per-point re-triangulation
 * CalcLSAXMulti (rpbacore.cpp:536-773, 777-837, 841-877).  Included by
 * rpba_oracle.c.  TEST INFRASTRUCTURE ONLY, like the rest of the oracle.
 *
 * The image partition is an input: the reference computes it with METIS
 * (rpbacore.cpp:945-1037), which is absent here; the tests hand the oracle the
 * partition the product computes, so both solve the same blocks.
 *
 * Plain FP64 sums where the reference sums in double (sig0g, the per-range
 * shares); the per-point costs, weight sums and point coordinates are float
 * where the reference keeps them in float (rpbacore.cpp:550-573, 600-615,
 * 797-834).  Parts run in OpenMP threads as in the reference; every result is
 * independent of the thread count (each part and each point range writes
 * only its own outputs, and the range sums are added in range order).
 */
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- blocks */

/* LSA in GCP mode for one part: LSA::LSA / InputLSAP + PointXXvIn with
 * gcpinfo (rbacore.cpp:37-123, 233-329, 1918-1974).  Images are the part's,
 * ascending; PM (12 each, double), f, d (float Camera values); points X
 * (3 each) with the sub-tracks toff/tcam (local image)/tfo (global file
 * index); gcpin per point (global GCP number or -1).  rect0: the constructor
 * path (the second outer iteration) re-orthonormalises image 0 before the
 * pass (rbacore.cpp:151-152); InputLSAP does it on the matrix it then
 * overwrites (rbacore.cpp:244-245, 253-254), so later iterations do not. */
static orc *orc_create_block(int nimg, const double *PM, const float *f, const float *d, int np, const double *X,
                             const int *toff, const int *tcam, const int *tfo, const float *obs_xy, const float *obs_w,
                             const int *gcpin, const int addf[3], float ihw2, int nrcam_unk, int rect0)
{
    orc *o = (orc *)xcalloc(1, sizeof(orc));
    const int nc = nimg;
    o->aoflag = 1;
    o->nrcam_unk = nrcam_unk;
    o->fsites = 0;
    o->solver = 0;
    o->robust = 0;
    for (int j = 0; j < 3; ++j) o->addf[j] = addf[j] != 0;
    o->nraddpi = o->addf[0] + o->addf[1] + o->addf[2];
    o->nimg = nc;
    o->XXnr = np;
    o->obs_xy = obs_xy;
    o->obs_w = obs_w;
    o->ihw2 = ihw2;
    o->defdiff = 5000. / sqrtf(ihw2);
    o->minimagenrdel = nc > 2 ? 3 : 2;
    /* PointXXvIn with GCP information (rbacore.cpp:1949-1962): a point stays
     * with two observations, or with one if it is a GCP -- tested as
     * gcpinfoin > 0, so GCP number 0 needs two */
    const int ntin = toff[np];
    o->nrpv = (int *)xcalloc((size_t)np + 1, sizeof(int));
    o->toff = (int *)xcalloc((size_t)np + 1, sizeof(int));
    o->timg = (int *)xcalloc((size_t)ntin + 1, sizeof(int));
    o->tfo = (int *)xcalloc((size_t)ntin + 1, sizeof(int));
    o->gcpi = (int *)xcalloc((size_t)np + 1, sizeof(int));
    int npk = 0, nt = 0;
    for (int i = 0; i < np; ++i) {
        const int t = toff[i + 1] - toff[i];
        if (fabs(X[3 * i + 2]) > 1.e-8 && (t >= 2 || (gcpin[i] > 0 && t > 0))) {
            o->nrpv[npk] = i;
            o->gcpi[npk] = gcpin[i];
            o->toff[npk] = nt;
            for (int q = toff[i]; q < toff[i + 1]; ++q) {
                o->timg[nt] = tcam[q];
                o->tfo[nt] = tfo[q];
                ++nt;
            }
            ++npk;
        }
    }
    o->toff[npk] = nt;
    o->npts = npk;
    o->ntr = nt;
    o->XX = (double *)xcalloc((size_t)3 * npk + 3, sizeof(double));
    o->XXn = (double *)xcalloc((size_t)3 * npk + 3, sizeof(double));
    for (int i = 0; i < npk; ++i)
        for (int j = 0; j < 3; ++j) o->XX[3 * i + j] = o->XXn[3 * i + j] = X[3 * o->nrpv[i] + j];
    o->PM = (double *)xcalloc((size_t)12 * nc, sizeof(double));
    memcpy(o->PM, PM, sizeof(double) * 12 * (size_t)nc);
    o->normcamera = -1; /* aoflag (rbacore.cpp:696-697) */
    o->camparv = (int *)xcalloc((size_t)nc + 2, sizeof(int));
    o->Rb = (double *)xcalloc((size_t)9 * nc, sizeof(double));
    o->Cb = (double *)xcalloc((size_t)3 * nc, sizeof(double));
    o->Rn = (double *)xcalloc((size_t)9 * nc, sizeof(double));
    o->Cn = (double *)xcalloc((size_t)3 * nc, sizeof(double));
    o->Rs = (double *)xcalloc((size_t)9 * nc, sizeof(double));
    o->Cs = (double *)xcalloc((size_t)3 * nc, sizeof(double));
    if (rect0) oracle_rectify(o->PM, 4);
    o->nrparir = 5 + o->nraddpi;
    o->nrpari = 6 + o->nraddpi;
    o->cf = (double *)xcalloc((size_t)nc, sizeof(double));
    o->cd0 = (double *)xcalloc((size_t)nc, sizeof(double));
    o->cd1 = (double *)xcalloc((size_t)nc, sizeof(double));
    o->nf = (double *)xcalloc((size_t)nc, sizeof(double));
    o->nd0 = (double *)xcalloc((size_t)nc, sizeof(double));
    o->nd1 = (double *)xcalloc((size_t)nc, sizeof(double));
    for (int i = 0; i < nc; ++i) {
        o->cf[i] = o->nf[i] = (double)f[i];
        o->cd0[i] = o->nd0[i] = (double)d[2 * i];
        o->cd1[i] = o->nd1[i] = (double)d[2 * i + 1];
    }
    o->lflflag = 1; /* sig0frac only matters past the first iteration, which itbreak never reaches */
    o->nrioff = o->nrpari * nc; /* rbacore.cpp:222-225 */
    /* XPMatrices holds 6 nrimages + 3 XXnr + 7 nrcameras (rbacore.cpp:118-120, 325-327), and
     * nrunknowns counts the intrinsics of all nrcameras = every image of the problem */
    o->xpsize = (size_t)nc * 6 + (size_t)3 * np + (size_t)7 * (size_t)(nrcam_unk > nc ? nrcam_unk : nc) + 16 + (size_t)nc * 3;
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

/* Adjust in GCP mode with itbreakflag and itmaxnum = 0, robust off
 * (rbacore.cpp:333-616 as CalcPLSA calls it, rpbacore.cpp:226): one
 * CalcBundleSep with multfact 1e-4 and the GCP priors, the trial decoded and
 * stored (the first iteration of a pass always is), then the return.  The
 * state it leaves is what OutputLSAP reads (PM; camerasnew = nf, nd0, nd1).
 * Returns 0 or the reference's error (nothing stored then).  pcg: the PCG
 * iterations. */
static int orc_block_adjust(orc *o, const float *W3, int nW3, const double *gcpc, int *pcg)
{
    o->W3 = W3;
    o->nW3 = nW3;
    o->gcpc = gcpc;
    *pcg = 0;
    const double mult = 1.e-4;
    o->imagepointnr = o->ntr;
    int nrgcp = 0;
    for (int i = 0; i < o->npts; ++i)
        if (o->gcpi[i] > -1) ++nrgcp;
    const int nrobsao = 3 * nrgcp;
    o->nrobs = 2 * o->imagepointnr + nrobsao;
    int rc = count_images(o, NULL, NULL); /* "No observations in image" goes to os, which CalcPLSA never prints */
    if (rc) return rc;
    memset(o->ISflag, 0, (size_t)o->ntr);
    memset(o->ISflagnew, 0, (size_t)o->ntr);
    prepare_weights(o);
    for (int i = 0; i < o->npts; ++i)
        for (int j = 0; j < 3; ++j) o->XP[o->nrioff + 3 * i + j] = o->XX[3 * i + j];
    encode_cameras(o);
    o->nrunknowns = o->nrioff3 + o->nraddpi * o->nrcam_unk; /* nrcameras = cset.size(), all images */
    memcpy(o->XXn, o->XX, sizeof(double) * 3 * (size_t)o->npts);
    memset(o->x, 0, sizeof(double) * o->xpsize);
    o->redundancy = o->nrobs - o->nrunknowns;
    if (o->redundancy < 0) return RPBA_E_LOW_REDUNDANCY;
    decode_cameras(o, o->XP);
    bundle_sep(o, mult);
    *pcg = o->pcg_last;
    for (int i = 0; i < o->nrunknowns; ++i) {
        if (isnan(o->x[i]) || isinf(o->x[i])) return RPBA_E_NAN_RESULT;
        o->XPn[i] = o->XP[i] - o->x[i];
    }
    decode_cameras(o, o->XPn);
    store(o); /* iter == 1: always stored (rbacore.cpp:565-570) */
    return 0;
}

/* ------------------------------------------------------- point kernel */

typedef struct {
    int N, P;
    const double *PMd;        /* PMatricesd, 12 per image */
    const float *cf, *cd;     /* cameraps: f, (d0, d1) per image (float, as Camera holds them) */
    const float *obs_xy, *obs_w;
    const int *part;          /* image -> part (imgindbv) */
    const float *sig0wv;      /* per image */
} lsax_in;

/* CalcCorrPointMulti (rpbacore.cpp:777-837): residuals, the float cost under
 * the current weights and the float key under the original ones, the float
 * normalised coordinates the derivatives use */
static float lsax_corr(const lsax_in *L, double X0, double X1, double X2, int n, const int *fo, const int *img,
                       const float *wx, const float *wy, const float *wxy, const float *wxo, const float *wyo,
                       const float *wxyo, double *corr, float *sig0v, float *p1v, float *p2v, float *p3v)
{
    float sig0 = 0.f;
    for (int js = 0; js < n; ++js) {
        const int j = img[js];
        const double *PM = L->PMd + 12 * (size_t)j;
        double p1 = PM[0] * X0 + PM[1] * X1 + PM[2] * X2 + PM[3];
        double p2 = PM[4] * X0 + PM[5] * X1 + PM[6] * X2 + PM[7];
        double p3 = PM[8] * X0 + PM[9] * X1 + PM[10] * X2 + PM[11];
        if (fabs(p3) < 1.e-3)
            p3 = (p3 > 0.) ? 1.e3 : -1.e3;
        else
            p3 = 1. / p3;
        p1 *= p3;
        p2 *= p3;
        p1v[js] = (float)p1;
        p2v[js] = (float)p2;
        p3v[js] = (float)p3;
        /* RadialReconstructionImage (rpba.h:850-857) with the float Camera */
        const double dist = p1 * p1 + p2 * p2;
        const double z = 1. + (double)((double)L->cd[2 * j] * dist + (double)L->cd[2 * j + 1] * dist * dist);
        p1 *= z;
        p2 *= z;
        p1 *= (double)L->cf[j];
        p2 *= (double)L->cf[j];
        const double dx = (double)L->obs_xy[2 * (size_t)fo[js]] - p1;
        const double dy = (double)L->obs_xy[2 * (size_t)fo[js] + 1] - p2;
        corr[2 * js] = dx;
        corr[2 * js + 1] = dy;
        const float si = (float)(dx * dx * wx[js] + dy * dy * wy[js] + 2 * dx * dy * wxy[js]);
        sig0 += si;
        sig0v[js] = (float)(dx * dx * wxo[js] + dy * dy * wyo[js] + 2 * dx * dy * wxyo[js]);
    }
    return sig0;
}

/* CalcLSAXMulti (rpbacore.cpp:536-773) for one point.  Track: n entries
 * fo[] (global file index), image from obs_cam.  Slots: the parts of the
 * point's initial track (ns of them, ascending part), pit[] = pitflags,
 * w3[6 ns] the stored W3 information per part.  X in/out.  Returns convflag;
 * *n_out = the track length after robust deletion (fo compacted in place),
 * 0 when the track is cleared.  sig0v_out (robustmode 1): the keys of the
 * entries, by position. */
static int lsax_point(const lsax_in *L, const int *obs_cam, double *X, int *fo, int n, int *n_out, int ns,
                      const int *spart, const unsigned char *pit, float *w3, int robustmode, int weightonly,
                      int wxyflag, double *sig0io, long *redundancy, float *sig0v_out)
{
    int iter = 1, iflag = robustmode > 0 ? 2 : 1;
    float sig0 = 1.e20f, sig0opt = 1.e20f;
    float X0 = (float)X[0], X1 = (float)X[1], X2 = (float)X[2];
    float Xold0 = X0, Xold1 = X1, Xold2 = X2;
    const float X0in = X0, X1in = X1, X2in = X2;
    int convflag = 0;
    const int pxs = n;
    int pxsd = pxs, pxs2 = pxs * 2;
    int *img = (int *)xcalloc((size_t)pxs + 1, sizeof(int));
    float *fb = (float *)xcalloc((size_t)pxs * 10 + 10, sizeof(float));
    float *p1v = fb, *p2v = fb + pxs, *p3v = fb + 2 * pxs, *wxi = fb + 3 * pxs, *wyi = fb + 4 * pxs, *wxyi = fb + 5 * pxs;
    float *wxio = fb + 6 * pxs, *wyio = fb + 7 * pxs, *wxyio = fb + 8 * pxs, *sig0v = fb + 9 * pxs;
    double *corrv = (double *)xcalloc((size_t)2 * pxs + 2, sizeof(double));
    double *Bw = (double *)xcalloc((size_t)6 * pxs + 6, sizeof(double)); /* eBpp2wv: rows 2js, 2js+1 */
    double *Bs = (double *)xcalloc((size_t)6 * pxs + 6, sizeof(double)); /* eBpp2 of each entry */
    unsigned char *delf = (unsigned char *)xcalloc((size_t)pxs + 1, 1);
    double *Nv = (double *)xcalloc((size_t)9 * ns + 9, sizeof(double)); /* eNppv per slot */
    float wsum = 0.f;
    for (int js = 0; js < pxs; ++js) {
        img[js] = obs_cam[fo[js]];
        const float *w = L->obs_w ? L->obs_w + 3 * (size_t)fo[js] : NULL;
        wxi[js] = wxio[js] = w ? w[0] : 1.f;
        wyi[js] = wyio[js] = w ? w[1] : 1.f;
        wxyi[js] = wxyio[js] = w ? w[2] : 0.f;
        wsum += wxi[js] + wyi[js];
    }
    sig0opt = lsax_corr(L, X0, X1, X2, pxs, fo, img, wxi, wyi, wxyi, wxio, wyio, wxyio, corrv, sig0v, p1v, p2v, p3v);
    if (weightonly) {
        wsum /= (float)pxs2;
        *sig0io += sig0opt / wsum;
        *redundancy += pxs2 - 3;
    }
    while (iflag > -1 && iter < 101) {
        double N[9] = {0};
        if (wxyflag) memset(Nv, 0, sizeof(double) * 9 * (size_t)ns);
        for (int js = 0; js < pxs; ++js) {
            const int j = img[js];
            const double *PM = L->PMd + 12 * (size_t)j;
            const double x = p1v[js], y = p2v[js];
            /* P1, P2: products of the float coordinates, rounded to float (rpbacore.cpp:617-619) */
            const double P3 = p3v[js], P1 = (double)(p1v[js] * p3v[js]), P2 = (double)(p2v[js] * p3v[js]);
            double B[2][3];
            for (int c = 0; c < 3; ++c) {
                B[0][c] = PM[c] * P3 - PM[8 + c] * P1;
                B[1][c] = PM[4 + c] * P3 - PM[8 + c] * P2;
            }
            /* CalcCalibratedDerivatives (rpbacore.cpp:841-877), skew 0 */
            {
                const double k2 = (double)L->cd[2 * j], k4 = (double)L->cd[2 * j + 1], fx = (double)L->cf[j];
                const double xy2 = x * x + y * y, xy4 = xy2 * xy2;
                const double xxpc = k4 * xy4 + x * (2 * k2 * x + 4 * k4 * x * xy2) + k2 * xy2 + 1.;
                const double xypc = x * (4 * k4 * xy2 * y + 2 * k2 * y);
                const double yxpc = y * (4 * k4 * x * xy2 + 2 * k2 * x);
                const double yypc = y * (2 * k2 * y + 4 * k4 * xy2 * y) + k4 * xy4 + k2 * xy2 + 1.;
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
            double *bw = Bw + 6 * (size_t)js;
            for (int c = 0; c < 3; ++c) {
                bw[c] = B[0][c] * wxi[js] + B[1][c] * wxyi[js];
                bw[3 + c] = B[1][c] * wyi[js] + B[0][c] * wxyi[js];
            }
            for (int c = 0; c < 3; ++c) {
                Bs[6 * (size_t)js + c] = B[0][c];
                Bs[6 * (size_t)js + 3 + c] = B[1][c];
            }
            double T[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) T[r * 3 + c] = B[0][r] * bw[c] + B[1][r] * bw[3 + c];
            for (int e = 0; e < 9; ++e) N[e] += T[e];
            if (wxyflag && L->P > 1) {
                const int ij = L->part[j];
                for (int s = 0; s < ns; ++s)
                    if (pit[s] && spart[s] != ij)
                        for (int e = 0; e < 9; ++e) Nv[9 * s + e] += T[e];
            }
        }
        const double m1 = 1.001;
        N[0] *= m1;
        N[4] *= m1;
        N[8] *= m1;
        double Ni[9];
        oracle_inv3(N, Ni);
        /* exip += (eNppi * Bw^T) * corr, the product taken left to right
         * (rpbacore.cpp:673-678) */
        double ex[3] = {0., 0., 0.};
        for (int i = 0; i < pxs; ++i) {
            const double *bw = Bw + 6 * (size_t)i;
            for (int r = 0; r < 3; ++r) {
                const double M0 = Ni[r * 3 + 0] * bw[0] + Ni[r * 3 + 1] * bw[1] + Ni[r * 3 + 2] * bw[2];
                const double M1 = Ni[r * 3 + 0] * bw[3] + Ni[r * 3 + 1] * bw[4] + Ni[r * 3 + 2] * bw[5];
                ex[r] += M0 * corrv[2 * i] + M1 * corrv[2 * i + 1];
            }
        }
        X0 = Xold0 + (float)ex[0];
        X1 = Xold1 + (float)ex[1];
        X2 = Xold2 + (float)ex[2];
        sig0 = lsax_corr(L, X0, X1, X2, pxs, fo, img, wxi, wyi, wxyi, wxio, wyio, wxyio, corrv, sig0v, p1v, p2v, p3v);
        if (iflag == 1 && robustmode == 2) {
            for (int js = 0; js < pxs; ++js) {
                if (sig0v[js] * L->sig0wv[img[js]] > 16.f) {
                    wxi[js] *= 1.e-4f;
                    wyi[js] *= 1.e-4f;
                    wxyi[js] *= 1.e-4f;
                    if (wxi[js] + wyi[js] < 2.e-8f) {
                        wxi[js] = wyi[js] = 1.e-8f;
                        wxyi[js] = 0.f;
                    }
                    --pxsd;
                    delf[js] = 1;
                }
            }
            if (pxsd < 2) break;
        }
        if ((double)sig0 < (double)sig0opt * 1.001) {
            if (sig0opt < sig0 * 1.05f) --iflag;
            if (!weightonly) {
                X[0] = Xold0 = X0;
                X[1] = Xold1 = X1;
                X[2] = Xold2 = X2;
            }
            sig0opt = sig0;
            convflag = 1;
            if (wxyflag)
                for (int s = 0; s < ns; ++s) {
                    const double *E = Nv + 9 * s;
                    float *w = w3 + 6 * s;
                    w[0] = (float)E[0]; /* XX */
                    w[1] = (float)E[4]; /* YY */
                    w[2] = (float)E[8]; /* ZZ */
                    w[3] = (float)E[1]; /* XY */
                    w[4] = (float)E[2]; /* XZ */
                    w[5] = (float)E[5]; /* YZ */
                }
        } else {
            --iflag;
        }
        if (weightonly) break;
        ++iter;
    }
    if (robustmode == 1 && sig0v_out)
        for (int js = 0; js < pxs; ++js) sig0v_out[js] = sig0v[js];
    int nk = pxs;
    if (robustmode == 2) {
        int ndel = 0, w = 0;
        for (int js = 0; js < pxs; ++js) {
            if (delf[js]) {
                ++ndel;
                continue;
            }
            fo[w++] = fo[js];
        }
        nk = w;
        pxs2 -= ndel * 2;
    }
    if (nk < 2 || !convflag) {
        nk = 0;
        convflag = 0;
        X[0] = X0in;
        X[1] = X1in;
        X[2] = X2in;
    } else if (!weightonly) {
        wsum /= (float)pxs2;
        *sig0io += sig0opt / wsum;
        *redundancy += pxs2 - 3;
    }
    *n_out = nk;
    free(img); free(fb); free(corrv); free(Bw); free(Bs); free(delf); free(Nv);
    return convflag;
}

/* -------------------------------------------------------- outer loop */

/* CalcPLSA's partitioned branch (rpbacore.cpp:100-510) over the block io
 * (the C ABI's flat arrays) with the image partition part[] (P parts, every
 * one non-empty).  Writes the outputs of the best outer iteration (pointXXvold,
 * PMatricesdold, XXdold, camerasold) into io.  Stats: itt = outer iterations
 * run, trace_* per outer iteration (sig0g, whether it was saved as the best,
 * the PCG iterations of its blocks, nrobs), sig0_final = the best sig0g.
 * fsites is not used (plain FP64 sums). */
int oracle_partition_solve(rpba_block *io, const rpba_flags *f, int P, const int *part, rpba_log_fn log, void *user,
                           rpba_stats *st)
{
    const int N = io->n_cams, pxs = io->n_pts, no = io->n_obs;
    if (P < 2 || N < P) return RPBA_E_INVALID;
    char buf[512];
    const int robust = f->robust != 0;
    const int addf0[3] = {f->est_k2 != 0, f->est_k4 != 0, f->est_f != 0};
    const int nraddp = addf0[0] + addf0[1] + addf0[2];
    float ihw2 = 0.f;
    for (int i = 0; i < N; ++i) {
        const int w = io->img_w ? io->img_w[i] : 1, h = io->img_h ? io->img_h[i] : 1;
        ihw2 += (float)((h + w) * (h + w)) * 0.25f;
    }
    ihw2 /= (float)N;
    /* parts: images ascending (imgindv), local index */
    int *pcnt = (int *)xcalloc((size_t)P + 1, sizeof(int));
    for (int i = 0; i < N; ++i) {
        if (part[i] < 0 || part[i] >= P) { free(pcnt); return RPBA_E_INVALID; }
        pcnt[part[i] + 1]++;
    }
    for (int p = 0; p < P; ++p) {
        if (pcnt[p + 1] == 0) { free(pcnt); return RPBA_E_INVALID; }
        pcnt[p + 1] += pcnt[p];
    }
    int *pimg = (int *)xcalloc((size_t)N, sizeof(int)), *lidx = (int *)xcalloc((size_t)N, sizeof(int));
    {
        int *pos = (int *)xcalloc((size_t)P, sizeof(int));
        memcpy(pos, pcnt, sizeof(int) * (size_t)P);
        for (int i = 0; i < N; ++i) {
            lidx[i] = pos[part[i]] - pcnt[part[i]];
            pimg[pos[part[i]]++] = i;
        }
        free(pos);
    }
    /* tracks in file order (InputBlock, rpba.cpp:219-238) */
    int64_t *base = (int64_t *)xcalloc((size_t)pxs + 1, sizeof(int64_t));
    int *len = (int *)xcalloc((size_t)pxs + 1, sizeof(int));
    for (int i = 0; i < no; ++i)
        if (io->obs_cam[i] >= 0) len[io->obs_pt[i]]++;
    for (int i = 0; i < pxs; ++i) base[i + 1] = base[i] + len[i];
    int *trk = (int *)xcalloc((size_t)base[pxs] + 1, sizeof(int));
    memset(len, 0, sizeof(int) * (size_t)pxs);
    for (int i = 0; i < no; ++i)
        if (io->obs_cam[i] >= 0) {
            const int p = io->obs_pt[i];
            trk[base[p] + len[p]++] = i;
        }
    /* per point: the parts of its initial track (slots), pitflags, W3 info */
    int64_t *soff = (int64_t *)xcalloc((size_t)pxs + 1, sizeof(int64_t));
    for (int i = 0; i < pxs; ++i) {
        int ns = 0;
        for (int q = 0; q < len[i]; ++q) {
            const int pp = part[io->obs_cam[trk[base[i] + q]]];
            int seen = 0;
            for (int r = 0; r < q; ++r) seen |= part[io->obs_cam[trk[base[i] + r]]] == pp;
            ns += !seen;
        }
        soff[i + 1] = soff[i] + ns;
    }
    int *spart = (int *)xcalloc((size_t)soff[pxs] + 1, sizeof(int));
    for (int i = 0; i < pxs; ++i) {
        int ns = 0;
        int64_t b = soff[i];
        for (int q = 0; q < len[i]; ++q) {
            const int pp = part[io->obs_cam[trk[base[i] + q]]];
            int s = 0;
            while (s < ns && spart[b + s] != pp) ++s;
            if (s == ns) {
                /* insert keeping the slots in ascending part order */
                int at = ns;
                while (at > 0 && spart[b + at - 1] > pp) { spart[b + at] = spart[b + at - 1]; --at; }
                spart[b + at] = pp;
                ++ns;
            }
        }
    }
    const int64_t NS = soff[pxs];
    unsigned char *pit = (unsigned char *)xcalloc((size_t)NS + 1, 1);
    unsigned char *bkept = (unsigned char *)xcalloc((size_t)NS + 1, 1);
    float *w3 = (float *)xcalloc((size_t)6 * NS + 6, sizeof(float));
    int *gcpinfo = (int *)xcalloc((size_t)pxs + 1, sizeof(int));
    for (int i = 0; i < pxs; ++i) gcpinfo[i] = -1;
    float *sig0wv = (float *)xcalloc((size_t)N, sizeof(float));
    /* state: XXd, PMatricesd, cameraps (= camerasnew), and the saved best */
    double *XXd = (double *)xcalloc((size_t)3 * pxs + 3, sizeof(double));
    for (int i = 0; i < 3 * pxs; ++i) XXd[i] = (double)io->pts[i];
    double *PMd = (double *)xcalloc((size_t)12 * N, sizeof(double));
    for (int i = 0; i < 12 * N; ++i) PMd[i] = (double)io->cam_P[i];
    float *cf = (float *)xcalloc((size_t)N, sizeof(float)), *cd = (float *)xcalloc((size_t)2 * N, sizeof(float));
    memcpy(cf, io->cam_f, sizeof(float) * (size_t)N);
    memcpy(cd, io->cam_d, sizeof(float) * 2 * (size_t)N);
    double *XXdold = (double *)xcalloc((size_t)3 * pxs + 3, sizeof(double));
    double *PMdold = (double *)xcalloc((size_t)12 * N, sizeof(double));
    memcpy(PMdold, PMd, sizeof(double) * 12 * (size_t)N);
    float *cfold = (float *)xcalloc((size_t)N, sizeof(float)), *cdold = (float *)xcalloc((size_t)2 * N, sizeof(float));
    int *trkold = (int *)xcalloc((size_t)base[pxs] + 1, sizeof(int));
    int *lenold = (int *)xcalloc((size_t)pxs + 1, sizeof(int));
    /* GCP centres and the per-part W3 lists of the next blocks */
    double *gcpc = (double *)xcalloc((size_t)3 * pxs + 3, sizeof(double));
    int *w3cnt = (int *)xcalloc((size_t)P + 1, sizeof(int));
    float **w3list = (float **)xcalloc((size_t)P, sizeof(float *));
    int *w3cap = (int *)xcalloc((size_t)P, sizeof(int));
    /* point ranges of the reference's threads (rpbacore.cpp:254-268) */
    const int nrppt = pxs / P;
    double *sig0gv = (double *)xcalloc((size_t)P, sizeof(double));
    long *redv = (long *)xcalloc((size_t)P, sizeof(long));
    float *keyv = (float *)xcalloc((size_t)no + 1, sizeof(float));   /* robust statistics pass: sig0v per file obs */
    unsigned char *keyset = (unsigned char *)xcalloc((size_t)no + 1, 1);
    int *bpcg = (int *)xcalloc((size_t)P, sizeof(int));
    const int nrparimages = (N - 2) * 6 + 5 + N * nraddp;
    double sig0opt = 1.e20;
    int iflag = robust ? 2 : 1, iter = 1, nit = 0;
    float sig0frac = 1.01f;
    double sig0g = 0.;
    if (st) { st->passes_done = 0; st->trace_len = 0; }

    lsax_in L;
    L.N = N; L.P = P; L.PMd = PMd; L.cf = cf; L.cd = cd; L.obs_xy = io->obs_xy; L.obs_w = io->obs_w; L.part = part;
    L.sig0wv = sig0wv;

    snprintf(buf, sizeof buf, "#threads: %d %d\n", P, (int)lroundf((float)N / 70.f));
    orc_log(log, user, buf);
    for (int p = 0; p < P; ++p) {
        snprintf(buf, sizeof buf, "Thread: %d #images %d indices %d %d\n", p, pcnt[p + 1] - pcnt[p], pcnt[p], pcnt[p + 1]);
        orc_log(log, user, buf);
    }

    while (iflag > 0 && iter < 51) {
        int robustit = 0, pcg_it = 0;
        if (iter > 1) {
            int addf[3] = {addf0[0], addf0[1], addf0[2]};
            if (iter < 3) addf[0] = addf[1] = addf[2] = 0;
            /* the blocks: each part's points (global order) with their
             * entries in that part (rpbacore.cpp:452-471, 208-237) */
            int64_t *lo = (int64_t *)xcalloc((size_t)P + 1, sizeof(int64_t));
            int64_t *eo = (int64_t *)xcalloc((size_t)P + 1, sizeof(int64_t));
            for (int i = 0; i < pxs; ++i)
                for (int q = 0; q < len[i]; ++q) {
                    const int pp = part[io->obs_cam[trk[base[i] + q]]];
                    int first = 1;
                    for (int r = 0; r < q; ++r) first &= part[io->obs_cam[trk[base[i] + r]]] != pp;
                    lo[pp + 1] += first;
                    eo[pp + 1] += 1;
                }
            for (int p = 0; p < P; ++p) { lo[p + 1] += lo[p]; eo[p + 1] += eo[p]; }
            int *plist = (int *)xcalloc((size_t)lo[P] + 1, sizeof(int));
            {
                int64_t *w = (int64_t *)xcalloc((size_t)P + 1, sizeof(int64_t));
                memcpy(w, lo, sizeof(int64_t) * (size_t)P);
                for (int i = 0; i < pxs; ++i)
                    for (int q = 0; q < len[i]; ++q) {
                        const int pp = part[io->obs_cam[trk[base[i] + q]]];
                        int first = 1;
                        for (int r = 0; r < q; ++r) first &= part[io->obs_cam[trk[base[i] + r]]] != pp;
                        if (first) plist[w[pp]++] = i;
                    }
                free(w);
            }
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < P; ++p) {
                const int nb = pcnt[p + 1] - pcnt[p];
                const int *im = pimg + pcnt[p];
                const int npp = (int)(lo[p + 1] - lo[p]);
                const int64_t ntp = eo[p + 1] - eo[p];
                const int *pl = plist + lo[p];
                int *gi = (int *)xcalloc((size_t)npp + 1, sizeof(int)), *tof = (int *)xcalloc((size_t)npp + 1, sizeof(int));
                int *tc = (int *)xcalloc((size_t)ntp + 1, sizeof(int)), *tf = (int *)xcalloc((size_t)ntp + 1, sizeof(int));
                int *gin = (int *)xcalloc((size_t)npp + 1, sizeof(int));
                double *Xb = (double *)xcalloc((size_t)3 * npp + 3, sizeof(double));
                int a = 0, e = 0;
                for (int ii = 0; ii < npp; ++ii) {
                    const int i = pl[ii];
                    int c = 0;
                    for (int q = 0; q < len[i]; ++q) {
                        const int fo = trk[base[i] + q];
                        if (part[io->obs_cam[fo]] != p) continue;
                        if (!c) { gi[a] = i; tof[a] = e; }
                        tc[e] = lidx[io->obs_cam[fo]];
                        tf[e] = fo;
                        ++e;
                        ++c;
                    }
                    if (c) {
                        gin[a] = gcpinfo[i];
                        for (int j = 0; j < 3; ++j) Xb[3 * a + j] = XXd[3 * i + j];
                        ++a;
                    }
                }
                tof[npp] = e;
                double *PMb = (double *)xcalloc((size_t)12 * nb, sizeof(double));
                float *fb = (float *)xcalloc((size_t)nb, sizeof(float)), *db = (float *)xcalloc((size_t)2 * nb, sizeof(float));
                for (int k = 0; k < nb; ++k) {
                    memcpy(PMb + 12 * k, PMd + 12 * (size_t)im[k], sizeof(double) * 12);
                    fb[k] = cf[im[k]];
                    db[2 * k] = cd[2 * im[k]];
                    db[2 * k + 1] = cd[2 * im[k] + 1];
                }
                orc *o = orc_create_block(nb, PMb, fb, db, npp, Xb, tof, tc, tf, io->obs_xy, io->obs_w, gin, addf, ihw2, N,
                                          iter == 2);
                int pc = 0;
                (void)orc_block_adjust(o, w3list[p], w3cnt[p], gcpc, &pc);
                bpcg[p] = pc;
                /* OutputLSAP (rbacore.cpp:620-640): the matrices, camerasnew as float */
                for (int k = 0; k < nb; ++k) {
                    memcpy(PMd + 12 * (size_t)im[k], o->PM + 12 * k, sizeof(double) * 12);
                    cf[im[k]] = (float)o->nf[k];
                    cd[2 * im[k]] = (float)o->nd0[k];
                    cd[2 * im[k] + 1] = (float)o->nd1[k];
                }
                /* kept points keep their entries in this part */
                int kk = 0;
                for (int j = 0; j < npp; ++j) {
                    const int i = gi[j];
                    const int kept = (kk < o->npts && o->nrpv[kk] == j);
                    if (kept) ++kk;
                    for (int64_t s = soff[i]; s < soff[i + 1]; ++s)
                        if (spart[s] == p) bkept[s] = (unsigned char)kept;
                }
                orc_destroy(o);
                free(gi); free(tof); free(tc); free(tf); free(gin); free(Xb); free(PMb); free(fb); free(db);
            }
            free(lo); free(eo); free(plist);
            for (int p = 0; p < P; ++p) pcg_it += bpcg[p];
        }
        if (iflag == 1) {
            robustit = robust;
            if (robust) sig0frac = 1.02f;
        }
        if (iter > 1) {
            /* pointXXv from the blocks' tracks, part by part (rpbacore.cpp:276-289) */
            int *tmp = (int *)xcalloc(64, sizeof(int));
            int tcap = 64;
            for (int i = 0; i < pxs; ++i) {
                if (len[i] == 0) continue;
                if (len[i] > tcap) { free(tmp); tcap = len[i]; tmp = (int *)xcalloc((size_t)tcap, sizeof(int)); }
                int w = 0;
                for (int64_t s = soff[i]; s < soff[i + 1]; ++s) {
                    const int p = spart[s];
                    int has = 0;
                    for (int q = 0; q < len[i]; ++q) has |= part[io->obs_cam[trk[base[i] + q]]] == p;
                    if (!has || !bkept[s]) continue;
                    for (int q = 0; q < len[i]; ++q) {
                        const int fo = trk[base[i] + q];
                        if (part[io->obs_cam[fo]] == p) tmp[w++] = fo;
                    }
                }
                memcpy(trk + base[i], tmp, sizeof(int) * (size_t)w);
                len[i] = w < 2 ? 0 : w;
            }
            free(tmp);
        }
        /* pitflags of the points that have entries (rpbacore.cpp:301-309) */
        for (int i = 0; i < pxs; ++i) {
            if (len[i] == 0) continue;
            for (int64_t s = soff[i]; s < soff[i + 1]; ++s) {
                int has = 0;
                for (int q = 0; q < len[i]; ++q) has |= part[io->obs_cam[trk[base[i] + q]]] == spart[s];
                pit[s] = (unsigned char)has;
            }
        }
        if (robustit) {
            /* robust statistics pass (rpbacore.cpp:311-350) */
            memset(keyset, 0, (size_t)no + 1);
#pragma omp parallel for schedule(static, 1)
            for (int t = 0; t < P; ++t) {
                const int b0 = nrppt * t, b1 = (t == P - 1) ? pxs : nrppt * (t + 1);
                double sg = 0.;
                long rd = 0;
                float *sv = NULL;
                int svc = 0;
                for (int i = b0; i < b1; ++i) {
                    if (len[i] == 0) continue;
                    if (len[i] > svc) { free(sv); svc = len[i]; sv = (float *)xcalloc((size_t)svc, sizeof(float)); }
                    int *fo = trk + base[i];
                    int fo0[1];
                    (void)fo0;
                    int nk = 0;
                    /* keys of the entries as they stand before the call */
                    int *fcopy = (int *)xcalloc((size_t)len[i], sizeof(int));
                    memcpy(fcopy, fo, sizeof(int) * (size_t)len[i]);
                    const int n0 = len[i];
                    lsax_point(&L, io->obs_cam, XXd + 3 * (size_t)i, fo, n0, &nk, (int)(soff[i + 1] - soff[i]),
                               spart + soff[i], pit + soff[i], w3 + 6 * soff[i], 1, 0, 0, &sg, &rd, sv);
                    for (int q = 0; q < n0; ++q) {
                        keyv[fcopy[q]] = sv[q];
                        keyset[fcopy[q]] = 1;
                    }
                    free(fcopy);
                    len[i] = nk;
                }
                free(sv);
            }
            /* sig0wv: 1 / (upper median * 2.1981) per image (rpbacore.cpp:357-375) */
            int *icnt = (int *)xcalloc((size_t)N + 1, sizeof(int));
            for (int q = 0; q < no; ++q)
                if (keyset[q]) icnt[io->obs_cam[q] + 1]++;
            for (int k = 0; k < N; ++k) icnt[k + 1] += icnt[k];
            float *kv = (float *)xcalloc((size_t)icnt[N] + 1, sizeof(float));
            int *pos = (int *)xcalloc((size_t)N + 1, sizeof(int));
            memcpy(pos, icnt, sizeof(int) * (size_t)N);
            for (int q = 0; q < no; ++q)
                if (keyset[q]) kv[pos[io->obs_cam[q]]++] = keyv[q];
            for (int k = 0; k < N; ++k) {
                const int n = icnt[k + 1] - icnt[k];
                if (n == 0) {
                    snprintf(buf, sizeof buf, "Image %d without observations!!!!!!!!!!!!!!!!!!!!!\n\n", k);
                    orc_log(log, user, buf);
                    sig0wv[k] = 0.f;
                } else {
                    sig0wv[k] = 1.f / (oracle_upper_median(kv + icnt[k], n) * 2.1981f);
                }
            }
            free(icnt); free(kv); free(pos);
        }
        const int robustmode = robustit ? 2 : 0, weightonly = iter == 1;
        /* the main point pass (rpbacore.cpp:378-392) */
#pragma omp parallel for schedule(static, 1)
        for (int t = 0; t < P; ++t) {
            const int b0 = nrppt * t, b1 = (t == P - 1) ? pxs : nrppt * (t + 1);
            double sg = 0.;
            long rd = 0;
            for (int i = b0; i < b1; ++i) {
                if (len[i] == 0) continue;
                int nk = 0;
                lsax_point(&L, io->obs_cam, XXd + 3 * (size_t)i, trk + base[i], len[i], &nk, (int)(soff[i + 1] - soff[i]),
                           spart + soff[i], pit + soff[i], w3 + 6 * soff[i], robustmode, weightonly, 1, &sg, &rd, NULL);
                len[i] = nk;
            }
            sig0gv[t] = sg;
            redv[t] = rd;
        }
        sig0g = 0.;
        long redundancy = 0;
        for (int t = 0; t < P; ++t) {
            sig0g += sig0gv[t];
            redundancy += redv[t];
        }
        int64_t nrobs = 0;
        for (int i = 0; i < pxs; ++i) nrobs += len[i];
        nrobs *= 2;
        sig0g /= (double)(redundancy - nrparimages);
        snprintf(buf, sizeof buf, "IT: %d %d  %g %g  %g %ld %lld %d %d  %ld\n", iter, iflag, sqrt(sig0g * (double)ihw2),
                 sig0opt / sig0g, sig0g, redundancy, (long long)nrobs, nrparimages, nraddp, redundancy - nrparimages);
        orc_log(log, user, buf);
        const int saved = sig0g <= sig0opt;
        if (saved) {
            memcpy(XXdold, XXd, sizeof(double) * 3 * (size_t)pxs);
            memcpy(PMdold, PMd, sizeof(double) * 12 * (size_t)N);
            memcpy(trkold, trk, sizeof(int) * (size_t)base[pxs]);
            memcpy(lenold, len, sizeof(int) * (size_t)pxs);
            memcpy(cfold, cf, sizeof(float) * (size_t)N);
            memcpy(cdold, cd, sizeof(float) * 2 * (size_t)N);
        }
        if (st && st->trace_sig0 && st->trace_len < st->trace_cap) {
            st->trace_sig0[st->trace_len] = sig0g;
            if (st->trace_accept) st->trace_accept[st->trace_len] = saved;
            if (st->trace_pcg) st->trace_pcg[st->trace_len] = pcg_it;
            if (st->trace_nrobs) st->trace_nrobs[st->trace_len] = (int)nrobs;
            st->trace_len++;
        }
        if (st) st->pcg_iters += pcg_it;
        /* GCPs: points seen by two parts or more (rpbacore.cpp:419-450) */
        {
            int nrgcp = 0;
            for (int p = 0; p < P; ++p) w3cnt[p] = 0;
            for (int i = 0; i < pxs; ++i) {
                if (len[i] < 2) continue;
                int gnum = 0;
                for (int64_t s = soff[i]; s < soff[i + 1] && gnum < 2; ++s) gnum += pit[s];
                if (gnum > 1) {
                    gcpinfo[i] = nrgcp;
                    for (int j = 0; j < 3; ++j) gcpc[3 * (size_t)nrgcp + j] = XXd[3 * i + j];
                    for (int64_t s = soff[i]; s < soff[i + 1]; ++s) {
                        if (!pit[s]) continue;
                        const int p = spart[s];
                        if (w3cnt[p] == w3cap[p]) {
                            const int nc2 = w3cap[p] ? 2 * w3cap[p] : 1024;
                            float *nl = (float *)xcalloc((size_t)6 * nc2, sizeof(float));
                            for (int v = 0; v < 6; ++v)
                                if (w3cap[p]) memcpy(nl + (size_t)v * nc2, w3list[p] + (size_t)v * w3cap[p], sizeof(float) * (size_t)w3cnt[p]);
                            free(w3list[p]);
                            w3list[p] = nl;
                            w3cap[p] = nc2;
                        }
                        for (int v = 0; v < 6; ++v) w3list[p][(size_t)v * w3cap[p] + w3cnt[p]] = w3[6 * s + v];
                        w3cnt[p]++;
                    }
                    ++nrgcp;
                } else {
                    gcpinfo[i] = -1;
                }
            }
            /* the lists are read with stride nW3 = their capacity */
            for (int p = 0; p < P; ++p)
                if (w3list[p] && w3cnt[p] < w3cap[p]) {
                    float *nl = (float *)xcalloc((size_t)6 * w3cnt[p] + 6, sizeof(float));
                    for (int v = 0; v < 6; ++v) memcpy(nl + (size_t)v * w3cnt[p], w3list[p] + (size_t)v * w3cap[p], sizeof(float) * (size_t)w3cnt[p]);
                    free(w3list[p]);
                    w3list[p] = nl;
                    w3cap[p] = w3cnt[p];
                }
        }
        /* iflag (rpbacore.cpp:479-493) */
        if (iter > 1) {
            if (sig0g <= sig0opt) {
                if (sig0opt < sig0g * (double)sig0frac && iter > 3)
                    --iflag;
                else
                    sig0opt = sig0g;
            } else if (iter > 3) {
                --iflag;
            }
        }
        ++iter;
        ++nit;
    }
    /* outputs of the saved state (rpbacore.cpp:497-507) */
    if (io->obs_kept) memset(io->obs_kept, 0, (size_t)no);
    if (io->pt_kept) memset(io->pt_kept, 0, (size_t)pxs);
    int64_t nobs_k = 0;
    int npts_k = 0;
    for (int i = 0; i < pxs; ++i) {
        for (int j = 0; j < 3; ++j) io->pts[3 * i + j] = (float)XXdold[3 * i + j];
        if (lenold[i] > 0) {
            ++npts_k;
            if (io->pt_kept) io->pt_kept[i] = 1;
        }
        for (int q = 0; q < lenold[i]; ++q) {
            if (io->obs_kept) io->obs_kept[trkold[base[i] + q]] = 1;
            ++nobs_k;
        }
    }
    for (int i = 0; i < 12 * N; ++i) io->cam_P[i] = (float)PMdold[i];
    memcpy(io->cam_f, cfold, sizeof(float) * (size_t)N);
    memcpy(io->cam_d, cdold, sizeof(float) * 2 * (size_t)N);
    if (st) {
        st->itt = nit;
        st->sig0_final = sig0opt;
        st->passes_done = 1;
        st->n_obs_kept = nobs_k;
        st->n_pts_kept = npts_k;
    }
    free(pcnt); free(pimg); free(lidx); free(base); free(len); free(trk); free(soff); free(spart); free(pit); free(bkept);
    free(w3); free(gcpinfo); free(sig0wv); free(XXd); free(PMd); free(cf); free(cd); free(XXdold); free(PMdold);
    free(cfold); free(cdold); free(trkold); free(lenold); free(gcpc); free(w3cnt); free(w3cap); free(sig0gv); free(redv);
    free(keyv); free(keyset); free(bpcg);
    for (int p = 0; p < P; ++p) free(w3list[p]);
    free(w3list);
    return 0;
}

/* test hook: CalcLSAXMulti for one point (lsax_point) on the given cameras;
 * part_of: the part of each of the nc images (P parts); slots spart/pit/w3 as
 * in lsax_point; sig0wv per image (robustmode 2).  out: [0] the sig0 share,
 * [1] the redundancy share, [2] the track length left. */
int oracle_lsax_point(const double *PMd, const float *cf, const float *cd, const float *obs_xy, const float *obs_w,
                      const int *obs_cam, int nc, const int *part_of, int P, const float *sig0wv, int *fo, int n,
                      double *X, int ns, const int *spart, const unsigned char *pit, float *w3, int robustmode,
                      int weightonly, int wxyflag, double *out)
{
    lsax_in L;
    L.N = nc; L.P = P; L.PMd = PMd; L.cf = cf; L.cd = cd; L.obs_xy = obs_xy; L.obs_w = obs_w; L.part = part_of;
    L.sig0wv = sig0wv;
    double sg = 0.;
    long rd = 0;
    int nk = 0;
    const int conv = lsax_point(&L, obs_cam, X, fo, n, &nk, ns, spart, pit, w3, robustmode, weightonly, wxyflag, &sg, &rd,
                                NULL);
    out[0] = sg;
    out[1] = (double)rd;
    out[2] = (double)nk;
    return conv;
}

/* test hooks: a part's LSA in GCP mode built from a plain block (its images
 * 0 .. n_cams - 1, tracks in the block's file order, gcpin per point), then
 * the pass set-up of the step up to the trial cameras of the linearisation,
 * so that the probes (orc_dims, orc_camparv, orc_get_theta, orc_residuals,
 * orc_jacobians, orc_linearize, orc_step) work on it */
orc *orc_create_block_io(const rpba_block *io, const rpba_flags *f, const int *gcpin, const float *W3, int nW3,
                         const double *gcpc, int nrcam_unk, int rect0)
{
    const int nc = io->n_cams, np = io->n_pts, no = io->n_obs;
    int *toff = (int *)xcalloc((size_t)np + 1, sizeof(int));
    for (int q = 0; q < no; ++q) toff[io->obs_pt[q] + 1]++;
    for (int i = 0; i < np; ++i) toff[i + 1] += toff[i];
    int *pos = (int *)xcalloc((size_t)np + 1, sizeof(int));
    memcpy(pos, toff, sizeof(int) * (size_t)np);
    int *tcam = (int *)xcalloc((size_t)no + 1, sizeof(int)), *tfo = (int *)xcalloc((size_t)no + 1, sizeof(int));
    for (int q = 0; q < no; ++q) {
        const int at = pos[io->obs_pt[q]]++;
        tcam[at] = io->obs_cam[q];
        tfo[at] = q;
    }
    double *PM = (double *)xcalloc((size_t)12 * nc, sizeof(double)), *X = (double *)xcalloc((size_t)3 * np + 3, sizeof(double));
    for (int i = 0; i < 12 * nc; ++i) PM[i] = (double)io->cam_P[i];
    for (int i = 0; i < 3 * np; ++i) X[i] = (double)io->pts[i];
    float ihw2 = 0.f;
    for (int i = 0; i < nc; ++i) {
        const int w = io->img_w ? io->img_w[i] : 1, h = io->img_h ? io->img_h[i] : 1;
        ihw2 += (float)((h + w) * (h + w)) * 0.25f;
    }
    ihw2 /= (float)nc;
    const int addf[3] = {f->est_k2 != 0, f->est_k4 != 0, f->est_f != 0};
    orc *o = orc_create_block(nc, PM, io->cam_f, io->cam_d, np, X, toff, tcam, tfo, io->obs_xy, io->obs_w, gcpin, addf,
                              ihw2, nrcam_unk, rect0);
    o->W3 = W3;
    o->nW3 = nW3;
    o->gcpc = gcpc;
    free(toff); free(pos); free(tcam); free(tfo); free(PM); free(X);
    return o;
}

int orc_setup_block(orc *o)
{
    o->imagepointnr = o->ntr;
    int nrgcp = 0;
    for (int i = 0; i < o->npts; ++i)
        if (o->gcpi[i] > -1) ++nrgcp;
    o->nrobs = 2 * o->ntr + 3 * nrgcp;
    const int rc = count_images(o, NULL, NULL);
    if (rc) return rc;
    prepare_weights(o);
    for (int i = 0; i < o->npts; ++i)
        for (int j = 0; j < 3; ++j) o->XP[o->nrioff + 3 * i + j] = o->XX[3 * i + j];
    encode_cameras(o);
    o->nrunknowns = o->nrioff3 + o->nraddpi * o->nrcam_unk;
    memcpy(o->XXn, o->XX, sizeof(double) * 3 * (size_t)o->npts);
    decode_cameras(o, o->XP);
    return 0;
}

/* the kept points' original indices (nrpv) and GCP numbers */
void orc_kept_points(const orc *o, int *nrpv, int *gcpi)
{
    for (int i = 0; i < o->npts; ++i) {
        nrpv[i] = o->nrpv[i];
        gcpi[i] = o->gcpi ? o->gcpi[i] : -1;
    }
}
