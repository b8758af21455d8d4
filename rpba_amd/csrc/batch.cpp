// Batched independent blocks: many bundle blocks solved at once, each with
// the global robust LM of LSA::Adjust (rbacore.cpp:333-616) and its own
// state, on one GPU.
//
// The blocks are joined into one union scene of block-diagonal structure
// (cameras, points and observations concatenated), so every kernel launch of
// an LM iteration covers all blocks: the linearisation, the Schur terms and
// the back-substitution are per entry / point / camera anyway, the reduced
// camera systems of the blocks are independent dense systems factored
// together by the batched MFMA Cholesky (dense.h), and the few block-level
// numbers (cost, weight sum, outlier count, NaN check) come from segmented
// reductions over each block's contiguous entry range.  The host keeps one
// LM state per block (sigma, iflag, iter, itt, accept); stores are masked by
// block.  Passes run in lockstep: a block that finishes a pass early sits out
// (its cameras are not decoded and nothing of it is stored) until every block
// has finished the pass, then RobustDelete and the next pass run for all.
// A block's result is therefore exactly what a solve of that block alone
// computes with the same kernels.
//
// The camera systems are solved either by the reference's block-Jacobi PCG
// (solver 0, SparsePCG of rbacore.cpp:1807-1914), segmented by block: one
// SpMV over the union S per iteration, each block with its own alpha, beta,
// best iterate, stopping rule and cap max(1000, nrioff), a block that stops
// leaving the rest running; or by the batched dense Cholesky (solver 1).
//
// The LM iteration's launch sequence is static between pass boundaries (the
// per-block inputs, iflag, the decode mask and the PCG caps, live in device
// memory), so it is captured once per pass in hipGraphs and replayed: one
// graph for the whole iteration with the dense solve; with the PCG one for
// the linearisation through the PCG set-up, one for four PCG iterations
// (replayed until every block has stopped; launches after a block stopped
// return at once) and one for the step, decode and trial cost.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/rpba.h"
#include "dense.h"
#include "host_util.h"
#include "kernels.h"
#include "partition.h"
#include "structure.h"
#include "structure_dev.h"

using namespace rpba;

namespace {

struct BlkState {
    int status = RPBA_OK;
    bool active = false;
    int iflag = 0, iter = 0, itt = 0;
    double sig0 = 0, sig0old = 0, sig0in = 0, wsum = 0, wsumg = 0, frac = 1.05;
    int imagepointnr = 0, nrobs = 0, nrunknowns = 0, redundancy = 0;
    int nrobsnew = 0, nrunknownsnew = 0, redundancynew = 0, nrunknownsgiven = 0, nrobsgiven = 0;
    int nc = 0, npts = 0, ntr = 0, nrioff = 0, normcamera = -1, esvmaxm = 0, lflflag = 0;
    rpba_stats stats{};
    std::vector<double> tr_sig0;
    std::vector<int> tr_acc, tr_nrobs, tr_pcg;
    int64_t pcg_total = 0;
};

constexpr int kSegChunk = 4;  // PCG iterations per replayed graph

}  // namespace

struct rpba_batch {
    int nb = 0;
    rpba_block *io = nullptr;
    rpba_flags fl{};
    rpba_batch_log_fn log = nullptr;
    void *user = nullptr;
    hipStream_t st = nullptr;
    int dev = 0;

    // union scene
    std::vector<int> u_cam, u_pt;
    std::vector<float> u_xy, u_w, u_P, u_f, u_d, u_pts;
    rpba_block U{};
    std::vector<int> coff, uptoff, uobsoff;  // block -> first camera / union point / union observation
    std::vector<int> camblk_h;
    int nc = 0, nraddpi = 0, addf[3] = {0, 0, 0}, min_keep = 3, rec32 = 0;
    float ihw2 = 1.f;
    double defdiff = 5000.;
    Structure s_init;
    std::vector<float> w0_init;
    std::vector<double> X_init, Pst, Ist, Pst_init, Ist_init, Rfirst, Cfirst;
    DevStruct ds;
    DenseSolver dense;
    std::vector<BlkState> B;
    int pass = 0, stage = 0;  // stage 0 = pass set-up due, 1 = iterating, 2 = done
    double mult = 1.e-3;
    bool structure_dirty = true;
    std::vector<int> eoff_h, poff_h;

    // device
    DBuf<float> d_wx, d_wy, d_wxy, d_key, d_s00;
    DBuf<unsigned char> d_flag, d_flagst, d_camflag, d_accept, d_Wrec, d_Yrec;
    DBuf<double> d_X, d_Xn, d_X0, d_Vinv, d_gp, d_Sd, d_b, d_sol, d_In_bak, d_Su, d_Dd, d_terms, d_bsum, d_partial;
    DBuf<double> d_bpart;
    DBuf<int> d_coff, d_uptoff, d_poff, d_eoff, d_camblk, d_ptblk, d_iflag, d_nanb, d_ipartial, d_iscal, d_npart;
    DBuf<double> d_scal;
    DBuf<double> c_Rb, c_Cb, c_Rn, c_Cn, c_In, c_th, c_thn, c_Pst, c_Ist;
    DBuf<unsigned> c_mask;
    DBuf<PcgScalars> d_pcg;
    double *h_bsum = nullptr;  // pinned: 3 per block
    int *h_nanb = nullptr;     // pinned: 1 per block

    // segmented PCG of the blocks' camera systems (solver 0)
    bool pcg = false;
    DBuf<double> d_Cinv, d_px, d_best, d_r, d_dv, d_Ad, d_h, d_h1, d_Sdf, d_ev, d_rowpart, d_colpart, d_dpart;
    DBuf<SegScal> d_segS;
    DBuf<int> d_ndone, d_live;
    int *h_ndone = nullptr;    // pinned
    SegScal *h_seg = nullptr;  // pinned: 1 per block
    int pcg_expect = 0;        // iterations the last lockstep PCG took (its slowest block)
    std::vector<int> roff_h;
    std::vector<long long> sblk_h;  // per block: its upper blocks of S (current structure)
    double spmv_bytes = 0.;    // algorithmic bytes of the SpMV launches timed (kernel timing on)

    // graph replay of the LM iteration: [0] the whole iteration (dense solve)
    // or the part before the PCG, [1] kSegChunk PCG iterations, [2] the part
    // after the PCG
    bool graph_on = true;
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    bool graph_valid = false;
    int64_t graph_launches = 0;

    // kernel timing (eager launches only)
    bool timing = false;
    std::vector<Timed> pending;
    std::map<std::string, std::pair<double, int64_t>> ktime;
    int iters_total = 0;

    CamPtrs cam() const
    {
        CamPtrs c;
        c.Rb = c_Rb.p; c.Cb = c_Cb.p; c.Rn = c_Rn.p; c.Cn = c_Cn.p; c.In = c_In.p;
        c.th = c_th.p; c.thn = c_thn.p; c.Pst = c_Pst.p; c.Ist = c_Ist.p; c.mask = c_mask.p;
        return c;
    }
    BatchArgs bargs() const
    {
        BatchArgs a;
        a.nb = nb; a.nc = nc; a.npts = ds.npts; a.ntr = ds.ntr;
        a.coff = d_coff.p; a.poff = d_poff.p; a.eoff = d_eoff.p;
        a.camblk = d_camblk.p; a.ptblk = d_ptblk.p; a.tpt = ds.tpt;
        return a;
    }
    void logb(int b, const char *line)
    {
        if (log) log(b, line, user);
    }
    template <class F>
    void timed(const char *name, F f)
    {
        if (!timing) {
            f();
            return;
        }
        Timed t{name, nullptr, nullptr, -1, false};
        (void)hipEventCreate(&t.a);
        (void)hipEventCreate(&t.b);
        (void)hipEventRecord(t.a, st);
        f();
        (void)hipEventRecord(t.b, st);
        pending.push_back(t);
    }
    void resolve_timing()
    {
        for (auto &t : pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
                auto &e = ktime[t.name];
                e.first += ms;
                e.second += 1;
            }
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        pending.clear();
    }
    void drop_graph()
    {
        for (int i = 0; i < 3; ++i) {
            if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
            if (graph[i]) (void)hipGraphDestroy(graph[i]);
            gexec[i] = nullptr;
            graph[i] = nullptr;
        }
        graph_valid = false;
    }
};

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------

// the union scene and the per-block constants of LSA::LSA, PrepareCameras and
// PrepareIteration (rbacore.cpp:37-229, 670-700)
static int batch_init(rpba_batch *c)
{
    const int nb = c->nb;
    c->coff.assign(nb + 1, 0);
    c->uptoff.assign(nb + 1, 0);
    c->uobsoff.assign(nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        const rpba_block &io = c->io[b];
        if (io.n_cams < 3 || io.n_pts < 0 || io.n_obs < 0 || !io.obs_cam || !io.obs_pt || !io.obs_xy || !io.cam_P ||
            !io.cam_f || !io.cam_d || !io.pts)
            return RPBA_E_INVALID;
        c->coff[b + 1] = c->coff[b] + io.n_cams;
        c->uptoff[b + 1] = c->uptoff[b] + io.n_pts;
        c->uobsoff[b + 1] = c->uobsoff[b] + io.n_obs;
    }
    const int nc = c->nc = c->coff[nb];
    const int np = c->uptoff[nb], no = c->uobsoff[nb];
    c->u_cam.resize(no);
    c->u_pt.resize(no);
    c->u_xy.resize(2 * (size_t)no);
    c->u_w.resize(3 * (size_t)no);
    c->u_P.resize(12 * (size_t)nc);
    c->u_f.resize(nc);
    c->u_d.resize(2 * (size_t)nc);
    c->u_pts.resize(3 * (size_t)np);
    c->camblk_h.resize(nc);
    float ihw2_0 = -1.f;
    for (int b = 0; b < nb; ++b) {
        const rpba_block &io = c->io[b];
        const int o0 = c->uobsoff[b], k0 = c->coff[b], p0 = c->uptoff[b];
        for (int q = 0; q < io.n_obs; ++q) {
            const int k = io.obs_cam[q];
            if (k >= io.n_cams || io.obs_pt[q] < 0 || io.obs_pt[q] >= io.n_pts) return RPBA_E_INVALID;
            c->u_cam[o0 + q] = k < 0 ? -1 : k + k0;
            c->u_pt[o0 + q] = io.obs_pt[q] + p0;
            c->u_xy[2 * (size_t)(o0 + q)] = io.obs_xy[2 * (size_t)q];
            c->u_xy[2 * (size_t)(o0 + q) + 1] = io.obs_xy[2 * (size_t)q + 1];
            for (int j = 0; j < 3; ++j)
                c->u_w[3 * (size_t)(o0 + q) + j] = io.obs_w ? io.obs_w[3 * (size_t)q + j] : (j < 2 ? 1.f : 0.f);
        }
        std::memcpy(c->u_P.data() + 12 * (size_t)k0, io.cam_P, sizeof(float) * 12 * (size_t)io.n_cams);
        std::memcpy(c->u_f.data() + k0, io.cam_f, sizeof(float) * (size_t)io.n_cams);
        std::memcpy(c->u_d.data() + 2 * (size_t)k0, io.cam_d, sizeof(float) * 2 * (size_t)io.n_cams);
        std::memcpy(c->u_pts.data() + 3 * (size_t)p0, io.pts, sizeof(float) * 3 * (size_t)io.n_pts);
        for (int k = 0; k < io.n_cams; ++k) c->camblk_h[k0 + k] = b;
        // ihw2 and the residual clamp (rpbacore.cpp:87-90, rbacore.cpp:48) are
        // kernel constants: one value for the batch
        float ihw2 = 0.f;
        for (int i = 0; i < io.n_cams; ++i) {
            const int w = io.img_w ? io.img_w[i] : 1, h = io.img_h ? io.img_h[i] : 1;
            ihw2 += (float)((h + w) * (h + w)) * 0.25f;
        }
        ihw2 /= (float)io.n_cams;
        if (b == 0) ihw2_0 = ihw2;
        if (ihw2 != ihw2_0) return RPBA_E_INVALID;
    }
    c->ihw2 = ihw2_0;
    c->defdiff = 5000. / (double)sqrtf(c->ihw2);
    c->U.n_cams = nc;
    c->U.n_pts = np;
    c->U.n_obs = no;
    c->U.obs_cam = c->u_cam.data();
    c->U.obs_pt = c->u_pt.data();
    c->U.obs_xy = c->u_xy.data();
    c->U.obs_w = c->u_w.data();
    c->U.cam_P = c->u_P.data();
    c->U.cam_f = c->u_f.data();
    c->U.cam_d = c->u_d.data();
    c->U.pts = c->u_pts.data();

    c->rec32 = c->fl.precision >= 1;
    c->addf[0] = c->fl.est_k2 != 0;
    c->addf[1] = c->fl.est_k4 != 0;
    c->addf[2] = c->fl.est_f != 0;
    c->nraddpi = c->addf[0] + c->addf[1] + c->addf[2];
    c->min_keep = 3;  // minimagenrdel for blocks of more than two images (rbacore.cpp:73-76)
    int rc = build_tracks(&c->U, c->s_init);
    if (rc) return rc;
    Structure &s = c->s_init;
    c->X_init.assign(3 * (size_t)s.npts, 0.);
    for (int i = 0; i < s.npts; ++i)
        for (int j = 0; j < 3; ++j) c->X_init[3 * (size_t)i + j] = (double)c->U.pts[3 * (size_t)s.nrpv[i] + j];
    c->w0_init.resize(3 * (size_t)s.ntr);
    for (int q = 0; q < s.ntr; ++q)
        for (int j = 0; j < 3; ++j) c->w0_init[3 * (size_t)q + j] = c->U.obs_w[3 * (size_t)s.tfo[q] + j];
    c->Pst.resize(12 * (size_t)nc);
    c->Ist.resize(3 * (size_t)nc);
    for (int k = 0; k < nc; ++k) {
        for (int j = 0; j < 12; ++j) c->Pst[12 * (size_t)k + j] = (double)c->U.cam_P[12 * (size_t)k + j];
        c->Ist[3 * (size_t)k] = (double)c->U.cam_d[2 * (size_t)k];
        c->Ist[3 * (size_t)k + 1] = (double)c->U.cam_d[2 * (size_t)k + 1];
        c->Ist[3 * (size_t)k + 2] = (double)c->U.cam_f[k];
    }
    c->B.assign(nb, BlkState());
    c->Rfirst.assign(9 * (size_t)nb, 0.);
    c->Cfirst.assign(3 * (size_t)nb, 0.);
    for (int b = 0; b < nb; ++b) {
        BlkState &S = c->B[b];
        const int k0 = c->coff[b], n = c->coff[b + 1] - k0;
        S.nc = n;
        // PrepareCameras: gauge camera from the unrectified matrices
        const double *P0 = c->Pst.data() + 12 * (size_t)k0;
        double T0[3];
        for (int r = 0; r < 3; ++r) T0[r] = P0[r * 4] * P0[3] + P0[r * 4 + 1] * P0[7] + P0[r * 4 + 2] * P0[11];
        double nd = 0.;
        S.normcamera = k0 + 1;
        for (int i = 1; i < n; ++i) {
            const double *P = c->Pst.data() + 12 * (size_t)(k0 + i);
            double d = 0.;
            for (int r = 0; r < 3; ++r) {
                const double Tv = P[r * 4] * P[3] + P[r * 4 + 1] * P[7] + P[r * 4 + 2] * P[11];
                d += (Tv - T0[r]) * (Tv - T0[r]);
            }
            if (d > nd) {
                nd = d;
                S.normcamera = k0 + i;
            }
        }
        // PrepareIteration: rectify the block's first image, its trial pose
        double *Pf = c->Pst.data() + 12 * (size_t)k0;
        rectify34(Pf);
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) c->Rfirst[9 * (size_t)b + r * 3 + cc] = Pf[r * 4 + cc];
        const double *R0 = c->Rfirst.data() + 9 * (size_t)b;
        for (int cc = 0; cc < 3; ++cc)
            c->Cfirst[3 * (size_t)b + cc] = R0[0 * 3 + cc] * Pf[3] + R0[1 * 3 + cc] * Pf[7] + R0[2 * 3 + cc] * Pf[11];
        S.lflflag = 0;
        for (int k = k0; k < k0 + n; ++k)
            if (c->U.cam_f[k] > 2.f) S.lflflag = 1;
        S.nrioff = c->nraddpi + (5 + c->nraddpi) + (6 + c->nraddpi) * (n - 2);
    }
    c->Pst_init = c->Pst;
    c->Ist_init = c->Ist;

    HIPCHK(hipGetDevice(&c->dev));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    if (hipHostMalloc((void **)&c->h_bsum, sizeof(double) * 3 * (size_t)nb, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_nanb, sizeof(int) * (size_t)nb, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_ndone, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_seg, sizeof(SegScal) * (size_t)nb, hipHostMallocDefault) != hipSuccess)
        return RPBA_E_NOMEM;
    const size_t n9 = 9 * (size_t)nc;
    rc = 0;
    rc |= c->d_Sd.alloc(81 * (size_t)nc);
    rc |= c->d_b.alloc(n9);
    rc |= c->d_sol.alloc(n9);
    rc |= c->d_s00.alloc((size_t)nc);
    rc |= c->d_scal.alloc(4);
    rc |= c->d_iscal.alloc(4);
    rc |= c->d_pcg.alloc(2);
    rc |= c->d_In_bak.alloc(3 * (size_t)nc);
    rc |= c->d_camflag.alloc((size_t)nc);
    rc |= c->d_accept.alloc((size_t)nb);
    rc |= c->d_iflag.alloc((size_t)nb);
    rc |= c->d_nanb.alloc((size_t)nb);
    rc |= c->d_bsum.alloc(3 * (size_t)nb);
    rc |= c->d_bpart.alloc(3 * (size_t)block_slices() * nb);
    rc |= c->d_npart.alloc((size_t)block_slices() * nb);
    rc |= c->d_coff.alloc((size_t)nb + 1);
    rc |= c->d_uptoff.alloc((size_t)nb + 1);
    rc |= c->d_poff.alloc((size_t)nb + 1);
    rc |= c->d_eoff.alloc((size_t)nb + 1);
    rc |= c->d_camblk.alloc((size_t)nc);
    rc |= c->c_Rb.alloc(n9);
    rc |= c->c_Cb.alloc(3 * (size_t)nc);
    if (c->pcg) {
        rc |= c->d_Cinv.alloc(81 * (size_t)nc);
        for (DBuf<double> *v : {&c->d_px, &c->d_best, &c->d_r, &c->d_dv, &c->d_Ad, &c->d_h, &c->d_h1, &c->d_Sdf})
            rc |= v->alloc(n9);
        rc |= c->d_ev.alloc((size_t)nc);
        rc |= c->d_segS.alloc((size_t)nb);
        rc |= c->d_ndone.alloc(1);
    }
    if (rc) return RPBA_E_NOMEM;
    rc |= c->d_coff.put(c->coff.data(), c->coff.size(), c->st);
    rc |= c->d_uptoff.put(c->uptoff.data(), c->uptoff.size(), c->st);
    rc |= c->d_camblk.put(c->camblk_h.data(), c->camblk_h.size(), c->st);
    if (rc) return RPBA_E_NOMEM;
    // the PCG's two scalar slots: slot 0's done flag gates the union SpMV
    PcgScalars z2[2];
    std::memset(z2, 0, sizeof z2);
    rc = c->d_pcg.put(z2, 2, c->st);
    if (rc) return rc;
    if (!c->pcg) {
        std::vector<int> ncam(nb);
        for (int b = 0; b < nb; ++b) ncam[b] = c->B[b].nc;
        rc = c->dense.plan(nb, c->coff.data(), ncam.data(), nc);
        if (rc) return rc;
        rc = c->dense.autotune(c->st);
        if (rc) return rc;
    }
    rc = c->ds.upload_tracks(c->s_init, c->U.obs_xy, c->w0_init, c->st);
    if (rc) return rc;
    rc |= c->d_X0.put(c->X_init.data(), c->X_init.size(), c->st);
    rc |= c->d_X.alloc(c->X_init.size());
    rc |= c->d_Xn.alloc(c->X_init.size());
    if (rc) return RPBA_E_NOMEM;
    HIPCHK(hipStreamSynchronize(c->st));
    return RPBA_OK;
}

static int restore_initial(rpba_batch *c)
{
    c->drop_graph();
    c->Pst = c->Pst_init;
    c->Ist = c->Ist_init;
    for (int b = 0; b < c->nb; ++b) {
        BlkState &S = c->B[b];
        const int keep_nc = S.nc, keep_nrioff = S.nrioff, keep_norm = S.normcamera, keep_lfl = S.lflflag;
        S = BlkState();
        S.nc = keep_nc;
        S.nrioff = keep_nrioff;
        S.normcamera = keep_norm;
        S.lflflag = keep_lfl;
        S.frac = S.lflflag ? (double)1.01f : (double)1.05f;
    }
    c->pass = 0;
    c->stage = 0;
    c->mult = 1.e-3;
    c->iters_total = 0;
    c->pcg_expect = 0;
    const size_t n = c->X_init.size();
    if (n) {
        HIPCHK(hipMemcpyAsync(c->d_X.p, c->d_X0.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_Xn.p, c->d_X0.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
    }
    // trial poses of the first images (never decoded) and trial intrinsics
    const int nc = c->nc;
    std::vector<double> Rn(9 * (size_t)nc, 0.), Cn(3 * (size_t)nc, 0.);
    for (int b = 0; b < c->nb; ++b) {
        const int k0 = c->coff[b];
        for (int j = 0; j < 9; ++j) Rn[9 * (size_t)k0 + j] = c->Rfirst[9 * (size_t)b + j];
        for (int j = 0; j < 3; ++j) Cn[3 * (size_t)k0 + j] = c->Cfirst[3 * (size_t)b + j];
    }
    int rc = 0;
    rc |= c->c_Rn.put(Rn.data(), Rn.size(), c->st);
    rc |= c->c_Cn.put(Cn.data(), Cn.size(), c->st);
    rc |= c->c_In.put(c->Ist.data(), c->Ist.size(), c->st);
    if (rc) return RPBA_E_NOMEM;
    HIPCHK(hipStreamSynchronize(c->st));
    c->structure_dirty = true;
    return RPBA_OK;
}

// camera-major order, pairs and blocks of the union, the per-block entry and
// point ranges, and the work arrays sized for them
static int build_structure(rpba_batch *c)
{
    int empty = -1;
    int rc = c->ds.build(c->st, &empty);
    if (rc) return rc;
    if (empty >= 0) {
        const int b = c->camblk_h[empty];
        if (c->B[b].nc > 3) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "BA-Error: No observations in image %d (image index)\n", empty - c->coff[b]);
            c->logb(b, buf);
        }
        return RPBA_E_NO_OBS_IN_IMAGE;
    }
    // kept points are in union point order, so each block's are contiguous:
    // its point and entry ranges and each point's block, on the device
    const int nb = c->nb, np = c->ds.npts;
    rc |= c->d_ptblk.alloc((size_t)std::max(np, 1));
    if (rc) return RPBA_E_NOMEM;
    launch_batch_ranges(np, nb, c->ds.nrpv, c->d_uptoff.p, c->ds.toff, c->d_ptblk.p, c->d_poff.p, c->d_eoff.p, c->st);
    HIPCHK(hipGetLastError());
    c->poff_h.assign(nb + 1, 0);
    c->eoff_h.assign(nb + 1, 0);
    HIPCHK(hipMemcpyAsync(c->poff_h.data(), c->d_poff.p, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(c->eoff_h.data(), c->d_eoff.p, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int j = 0; j < nb; ++j) {
        c->B[j].npts = c->poff_h[j + 1] - c->poff_h[j];
        c->B[j].ntr = c->eoff_h[j + 1] - c->eoff_h[j];
    }
    const size_t ntr = (size_t)c->ds.ntr;
    rc |= c->d_wx.alloc(ntr);
    rc |= c->d_wy.alloc(ntr);
    rc |= c->d_wxy.alloc(ntr);
    rc |= c->d_key.alloc(ntr);
    rc |= c->d_flag.alloc(ntr);
    rc |= c->d_flagst.alloc(ntr);
    rc |= c->d_terms.alloc(ntr);
    const size_t esz = c->rec32 ? sizeof(float) : sizeof(double);
    rc |= c->d_Wrec.alloc((size_t)kRec * esz * ntr);
    rc |= c->d_Yrec.alloc((size_t)kRec * esz * ntr);
    rc |= c->d_Vinv.alloc(9 * (size_t)np);
    rc |= c->d_gp.alloc(3 * (size_t)np);
    rc |= c->d_Su.alloc(81 * (size_t)std::max(c->ds.nblk, 1));
    rc |= c->d_Dd.alloc(90 * (size_t)std::max(c->ds.nch, 1));
    const int nbk = std::max(cost_blocks(c->ds.ntr), (int)((np + 255) / 256 + 1));
    rc |= c->d_partial.alloc((size_t)nbk);
    rc |= c->d_ipartial.alloc((size_t)nbk);
    if (c->pcg) {
        rc |= c->d_rowpart.alloc(63 * (size_t)std::max(c->ds.nchunk, 1));
        rc |= c->d_colpart.alloc(9 * (size_t)std::max(c->ds.nslab, 1));
        rc |= c->d_dpart.alloc((size_t)std::max(spmv_grid(c->ds.nblk), 1));
        rc |= c->d_live.alloc((size_t)std::max(c->ds.nstrip, 1));
    }
    if (rc) return RPBA_E_NOMEM;
    if (c->pcg) {
        // the upper blocks are in row order: block b holds those of rows
        // coff[b] .. coff[b+1] - 1
        c->roff_h.assign((size_t)c->nc + 1, 0);
        HIPCHK(hipMemcpyAsync(c->roff_h.data(), c->ds.roff, sizeof(int) * c->roff_h.size(), hipMemcpyDeviceToHost,
                              c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->sblk_h.assign(c->nb, 0);
        for (int b = 0; b < c->nb; ++b) c->sblk_h[b] = c->roff_h[c->coff[b + 1]] - c->roff_h[c->coff[b]];
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->structure_dirty = false;
    return RPBA_OK;
}

// EncodeCameras (rbacore.cpp:999-1087) for every block: the first image of a
// block keeps its pose, the block's gauge camera fixes its C component of
// largest magnitude
static int encode_cameras(rpba_batch *c)
{
    const int nc = c->nc;
    std::vector<double> Rb(9 * (size_t)nc), Cb(3 * (size_t)nc), th(9 * (size_t)nc, 0.);
    std::vector<unsigned> mask((size_t)nc, 0);
    for (int k = 0; k < nc; ++k) {
        const int b = c->camblk_h[k];
        const bool first = k == c->coff[b];
        double *P = c->Pst.data() + 12 * (size_t)k;
        if (!first) rectify34(P);
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) Rb[9 * (size_t)k + r * 3 + cc] = P[r * 4 + cc];
        const double *R = Rb.data() + 9 * (size_t)k;
        for (int cc = 0; cc < 3; ++cc)
            Cb[3 * (size_t)k + cc] = R[0 * 3 + cc] * P[3] + R[1 * 3 + cc] * P[7] + R[2 * 3 + cc] * P[11];
        unsigned m = 0;
        if (!first) m |= 0x3fu;
        if (c->addf[0]) m |= 1u << 6;
        if (c->addf[1]) m |= 1u << 7;
        if (c->addf[2]) m |= 1u << 8;
        if (!first && k == c->B[b].normcamera) {
            const double *C = Cb.data() + 3 * (size_t)k;
            float maxv = 0.f;
            int maxi = 0;
            for (int i = 0; i < 3; ++i)
                if (std::fabs(C[i]) > maxv) {
                    maxv = (float)std::fabs(C[i]);
                    maxi = i;
                }
            c->B[b].esvmaxm = maxi;
            m &= ~(1u << (3 + maxi));
        }
        mask[k] = m;
        for (int j = 0; j < 3; ++j) th[9 * (size_t)k + 6 + j] = c->Ist[3 * (size_t)k + j];
    }
    int rc = 0;
    rc |= c->c_Rb.put(Rb.data(), Rb.size(), c->st);
    rc |= c->c_Cb.put(Cb.data(), Cb.size(), c->st);
    rc |= c->c_th.put(th.data(), th.size(), c->st);
    rc |= c->c_thn.put(th.data(), th.size(), c->st);
    rc |= c->c_mask.put(mask.data(), mask.size(), c->st);
    rc |= c->c_Pst.put(c->Pst.data(), c->Pst.size(), c->st);
    rc |= c->c_Ist.put(c->Ist.data(), c->Ist.size(), c->st);
    if (rc) return RPBA_E_NOMEM;
    HIPCHK(hipStreamSynchronize(c->st));
    return RPBA_OK;
}

static LinArgs lin_args(rpba_batch *c)
{
    LinArgs a;
    a.nc = c->nc; a.npts = c->ds.npts; a.ntr = c->ds.ntr; a.npgrp = c->ds.npgrp; a.pgrp = c->ds.pgrp; a.pg4 = c->ds.pg4;
    a.defdiff = c->defdiff; a.mult1 = 1. + c->mult;
    a.cam = c->cam();
    a.toff = c->ds.toff; a.tcam = c->ds.tcam; a.tpt = c->ds.tpt; a.camoff = c->ds.camoff; a.camq = c->ds.camq;
    a.ox = c->ds.ox; a.oy = c->ds.oy; a.wx = c->d_wx.p; a.wy = c->d_wy.p; a.wxy = c->d_wxy.p;
    a.X = c->d_X.p;
    a.Wrec = c->d_Wrec.p; a.Yrec = c->d_Yrec.p; a.rec32 = c->rec32;
    a.Vinv = c->d_Vinv.p; a.gp = c->d_gp.p; a.Sd = c->d_Sd.p; a.bvec = c->d_b.p;
    a.Dd = c->d_Dd.p; a.nch = c->ds.nch; a.ch_start = c->ds.ch_start; a.ch_off = c->ds.ch_off;
    a.Dk = nullptr;
    a.yrank = nullptr;
    return a;
}

static StepArgs step_args(rpba_batch *c, const double *sol)
{
    StepArgs a;
    a.nc = c->nc; a.npts = c->ds.npts; a.cam = c->cam();
    a.toff = c->ds.toff; a.tcam = c->ds.tcam; a.pgrp = c->ds.pgrp; a.npgrp = c->ds.npgrp; a.pg4 = c->ds.pg4;
    a.Wrec = c->d_Wrec.p; a.rec32 = c->rec32;
    a.Vinv = c->d_Vinv.p; a.gp = c->d_gp.p; a.sol = sol; a.X = c->d_X.p; a.Xn = c->d_Xn.p;
    a.S = c->d_pcg.p;
    a.camflag = c->d_camflag.p;
    return a;
}

static CostArgs cost_args(rpba_batch *c)
{
    CostArgs a;
    a.ntr = c->ds.ntr; a.robust = c->fl.robust != 0; a.write_key = 1; a.defdiff = c->defdiff;
    a.cam = c->cam(); a.tcam = c->ds.tcam; a.tpt = c->ds.tpt;
    a.ox = c->ds.ox; a.oy = c->ds.oy; a.wx = c->d_wx.p; a.wy = c->d_wy.p; a.wxy = c->d_wxy.p;
    a.w0x = c->ds.w0x; a.w0y = c->ds.w0y; a.w0xy = c->ds.w0xy;
    a.X = c->d_Xn.p; a.key = c->d_key.p; a.partial = c->d_partial.p;
    a.terms = c->d_terms.p;
    return a;
}

// the union SpMV of the PCG (S formed by strip slabs) and its per-block
// segmentation
static PcgArgs seg_pcg_args(rpba_batch *c)
{
    PcgArgs pa;
    std::memset(&pa, 0, sizeof pa);
    const DevStruct &d = c->ds;
    pa.nc = c->nc; pa.msize = 1000; pa.Sd = c->d_Sd.p; pa.Su = c->d_Su.p; pa.Cinv = c->d_Cinv.p; pa.b = c->d_b.p;
    pa.row_off = d.roff; pa.col_off = d.coff; pa.col_blk = d.cblk; pa.blk_k = d.bk; pa.blk_l = d.bl;
    pa.x = c->d_px.p; pa.best = c->d_best.p; pa.r = c->d_r.p; pa.d = c->d_dv.p; pa.Ad = c->d_Ad.p;
    pa.h = c->d_h.p; pa.h1 = c->d_h1.p; pa.part = nullptr;
    pa.sm = d.sumap(); pa.nslab = d.nslab; pa.nchunk = d.nchunk; pa.nstrip = d.nstrip; pa.slab_chunk = d.slab_chunk;
    pa.slab_bs = d.slab_bs; pa.slab_l = d.slab_l; pa.slab_mask = d.slab_mask; pa.strip_s0 = d.strip_s0;
    pa.chunk_off = d.chunk_off; pa.chunk_strip = d.chunk_strip; pa.cs_off = d.cs_off; pa.cs_list = d.cs_list;
    pa.rowpart = c->d_rowpart.p; pa.colpart = c->d_colpart.p; pa.dpart = c->d_dpart.p; pa.nblk = d.nblk;
    pa.ndpart = spmv_grid(d.nblk); pa.dsum = c->d_dpart.p; pa.ndsum = pa.ndpart; pa.S = c->d_pcg.p; pa.j = 0;
    pa.implicit = 0; pa.npts = d.npts; pa.ntr = d.ntr; pa.nch = d.nch;
    pa.strip_live = c->d_live.p;
    return pa;
}

static SegArgs seg_args(rpba_batch *c)
{
    SegArgs sg;
    sg.nc = c->nc; sg.nb = c->nb; sg.boff = c->d_coff.p; sg.cblk = c->d_camblk.p; sg.S = c->d_segS.p;
    sg.ndone = c->d_ndone.p; sg.Sdf = c->d_Sdf.p; sg.e = c->d_ev.p;
    sg.strip_live = c->d_live.p; sg.nstrip = c->ds.nstrip;
    return sg;
}

// the launches of one LM iteration before the camera solve's iterations
// (reweighting in pass 0 of a robust solve, CalcBundleSep), then the dense
// solve, or the block-Jacobi inverses and the PCG set-up
static void launches_pre(rpba_batch *c)
{
    const bool rw = c->pass == 0 && c->fl.robust;
    if (rw) {
        RobustArgs ra;
        ra.nc = c->nc; ra.ntr = c->ds.ntr; ra.iflag = 0;
        ra.camoff = c->ds.camoff; ra.camq = c->ds.camq; ra.tpt = c->ds.tpt; ra.toff = c->ds.toff; ra.tcam = c->ds.tcam;
        ra.key = c->d_key.p; ra.w0x = c->ds.w0x; ra.w0y = c->ds.w0y; ra.w0xy = c->ds.w0xy;
        ra.s00 = c->d_s00.p; ra.wx = c->d_wx.p; ra.wy = c->d_wy.p; ra.wxy = c->d_wxy.p;
        ra.flag = c->d_flag.p; ra.partial = c->d_partial.p; ra.ipartial = c->d_ipartial.p;
        ra.iflag_blk = c->d_iflag.p;
        ra.camblk = c->d_camblk.p;
        c->timed("reweight", [&] { launch_reweight(ra, c->d_scal.p + 1, c->d_iscal.p, c->st); });
    }
    LinArgs la = lin_args(c);
    SchurArgs sa;
    sa.nblk = c->ds.nblk; sa.poff = c->ds.poff; sa.pab = c->ds.pab;
    sa.Yrec = c->d_Yrec.p; sa.Wrec = c->d_Wrec.p; sa.Su = c->d_Su.p; sa.sm = c->ds.sumap(); sa.sperm = c->ds.sperm; sa.nslab = c->ds.nslab; sa.slab_bs = c->ds.slab_bs; sa.npairs = c->ds.npairs; sa.rec32 = c->rec32;
    c->timed("lin_points", [&] { launch_lin_points(la, c->st); });
    c->timed("schur_pairs", [&] { launch_schur_pairs(sa, c->st); });
    c->timed("schur_diag", [&] { launch_schur_diag(la, c->st); });
    c->timed("lin_cameras", [&] { launch_lin_cameras(la, c->st); });
    if (c->pcg) {
        c->timed("bj_inverse", [&] { launch_bj_inverse(c->nc, c->d_Sd.p, c->d_Cinv.p, c->st); });
        const PcgArgs pa = seg_pcg_args(c);
        const SegArgs sg = seg_args(c);
        c->timed("pcg_init", [&] { launch_seg_init(pa, sg, c->st); });
        return;
    }
    c->timed("dense_assemble", [&] {
        c->dense.assemble(c->d_Sd.p, c->d_Su.p, c->ds.sumap(), c->ds.bk, c->ds.bl, c->ds.nblk, c->d_b.p, c->st);
    });
    c->dense.factor_solve(c->d_sol.p, c->st, [&](const char *nm, auto f) { c->timed(nm, f); });
}

// n iterations of the segmented PCG
static void launches_pcg(rpba_batch *c, int n)
{
    const PcgArgs pa = seg_pcg_args(c);
    const SegArgs sg = seg_args(c);
    for (int i = 0; i < n; ++i) {
        c->timed("spmv", [&] { launch_spmv(pa, c->st); });
        c->timed("seg_update", [&] { launch_seg_update(pa, sg, c->st); });
    }
}

// the rest of the iteration: back-substitution, decode, trial cost, the
// per-block sums and NaN flags
static void launches_post(rpba_batch *c)
{
    const bool rw = c->pass == 0 && c->fl.robust;
    const double *sol = c->pcg ? c->d_best.p : c->d_sol.p;
    StepArgs sa2 = step_args(c, sol);
    c->timed("backsub", [&] { launch_backsub(sa2, c->st); });
    launch_copy(c->d_In_bak.p, c->c_In.p, 3 * (size_t)c->nc, c->st);
    c->timed("decode", [&] { launch_decode(sa2, c->st); });
    CostArgs ca = cost_args(c);
    c->timed("cost", [&] { launch_cost(ca, c->d_scal.p, c->st); });
    BatchArgs ba = c->bargs();
    c->timed("block_sums", [&] {
        launch_block_sums(ba, c->d_terms.p, rw ? c->d_wx.p : nullptr, rw ? c->d_wy.p : nullptr,
                          rw ? c->d_flag.p : nullptr, c->d_bsum.p, c->d_bpart.p, c->st);
        launch_block_nan(ba, sol, c->d_Xn.p, c->d_nanb.p, c->d_npart.p, c->st);
    });
}

static int capture_one(rpba_batch *c, int i, void (*f)(rpba_batch *))
{
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    f(c);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(c->st, &g));
    c->graph[i] = g;
    HIPCHK(hipGraphInstantiate(&c->gexec[i], g, nullptr, nullptr, 0));
    return RPBA_OK;
}

static int capture_graph(rpba_batch *c)
{
    c->drop_graph();
    int rc;
    if (!c->pcg) {
        rc = capture_one(c, 0, [](rpba_batch *cc) {
            launches_pre(cc);
            launches_post(cc);
        });
    } else {
        rc = capture_one(c, 0, launches_pre);
        if (!rc) rc = capture_one(c, 1, [](rpba_batch *cc) { launches_pcg(cc, kSegChunk); });
        if (!rc) rc = capture_one(c, 2, launches_post);
    }
    if (rc) return rc;
    c->graph_valid = true;
    return RPBA_OK;
}

// the segmented PCG until every block has stopped: as many iterations as the
// last lockstep solve took before the first look at the count, then
// kSegChunk at a time.  With kernel timing on, one iteration at a time, so
// that no launch after the last block stopped enters the timings.
static int run_pcg(rpba_batch *c, bool replay)
{
    int cap = 0;
    for (int b = 0; b < c->nb; ++b) cap = std::max(cap, c->h_seg[b].msize);
    const int per = c->timing ? 1 : kSegChunk;
    int n = c->timing ? 1 : std::max(1, (c->pcg_expect + kSegChunk - 1) / kSegChunk), launched = 0;
    for (;;) {
        for (int i = 0; i < n; ++i) {
            if (replay)
                HIPCHK(hipGraphLaunch(c->gexec[1], c->st));
            else
                launches_pcg(c, per);
        }
        launched += n * per;
        HIPCHK(hipMemcpyAsync(c->h_ndone, c->d_ndone.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (*c->h_ndone >= c->nb) return RPBA_OK;
        if (launched > cap + kSegChunk) return RPBA_E_HIP;  // cannot happen: every block stops at its cap
        n = 1;
    }
}

// per-block sums of the trial cost, readback; returns 0
static int block_costs(rpba_batch *c, bool with_weights)
{
    CostArgs ca = cost_args(c);
    launch_cost(ca, c->d_scal.p, c->st);
    BatchArgs ba = c->bargs();
    launch_block_sums(ba, c->d_terms.p, with_weights ? c->ds.w0x : nullptr, with_weights ? c->ds.w0y : nullptr,
                      nullptr, c->d_bsum.p, c->d_bpart.p, c->st);
    HIPCHK(hipMemcpyAsync(c->h_bsum, c->d_bsum.p, sizeof(double) * 3 * (size_t)c->nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return RPBA_OK;
}

// Adjust's pass set-up (rbacore.cpp:365-508) for every block
static int pass_setup(rpba_batch *c)
{
    char buf[160];
    c->mult = c->pass > 0 ? 1.e-6 : 1.e-3;
    if (c->structure_dirty) {
        const int rc = build_structure(c);
        if (rc) return rc;
    }
    const size_t ntr = (size_t)c->ds.ntr;
    // PrepareWeights2D: current weights = original ones, flags cleared
    if (ntr) {
        HIPCHK(hipMemcpyAsync(c->d_wx.p, c->ds.w0x, sizeof(float) * ntr, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_wy.p, c->ds.w0y, sizeof(float) * ntr, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_wxy.p, c->ds.w0xy, sizeof(float) * ntr, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemsetAsync(c->d_flag.p, 0, ntr, c->st));
        HIPCHK(hipMemsetAsync(c->d_flagst.p, 0, ntr, c->st));
    }
    int rc = encode_cameras(c);
    if (rc) return rc;
    if (c->pass == 0) {
        // decode theta itself for every block, trial points = points
        std::vector<unsigned char> cf(c->nc);
        for (int k = 0; k < c->nc; ++k) cf[k] = (unsigned char)(1 | (k == c->coff[c->camblk_h[k]] ? 2 : 0));
        rc = c->d_camflag.put(cf.data(), cf.size(), c->st);
        if (rc) return rc;
        StepArgs sa = step_args(c, nullptr);
        launch_decode(sa, c->st);
        HIPCHK(hipMemcpyAsync(c->d_Xn.p, c->d_X.p, sizeof(double) * 3 * (size_t)c->ds.npts, hipMemcpyDeviceToDevice, c->st));
    }
    rc = block_costs(c, true);  // sum of w0x + w0y per block, and (pass 0) the initial costs
    if (rc) return rc;
    for (int b = 0; b < c->nb; ++b) {
        BlkState &S = c->B[b];
        if (S.status != RPBA_OK) continue;
        S.imagepointnr = S.ntr;
        S.nrobs = 2 * S.imagepointnr;
        S.sig0old = (double)1.e20f;
        S.wsum = (double)(2 * S.imagepointnr) / c->h_bsum[3 * b + 1];
        S.wsumg = S.wsum;
        const int nrioff3 = S.nrioff + 3 * S.npts;
        S.nrunknowns = nrioff3 + c->nraddpi * S.nc;  // counts the intrinsics twice (rbacore.cpp:462)
        S.redundancy = S.nrobs - S.nrunknowns;
        if (S.redundancy < 0) {
            if (S.nc > 3) c->logb(b, "BA-Error: Low redundancy\n");
            S.status = RPBA_E_LOW_REDUNDANCY;
            continue;
        }
        if (c->pass == 0) {
            S.sig0 = c->h_bsum[3 * b] * S.wsumg / (double)S.redundancy;
            S.sig0in = S.sig0old = S.sig0;
        } else {
            S.sig0in = S.sig0old;
        }
        S.iflag = c->fl.robust ? 2 : 1;
        S.iter = 1;
        S.nrunknownsgiven = S.nrunknowns;
        S.nrobsgiven = S.nrobs;
        S.active = true;
    }
    (void)buf;
    c->graph_valid = false;
    c->stage = 1;
    return RPBA_OK;
}

static int pass_end(rpba_batch *c)
{
    char buf[256];
    for (int b = 0; b < c->nb; ++b) {
        BlkState &S = c->B[b];
        if (S.status != RPBA_OK) continue;
        const double s_in = std::sqrt(S.sig0in * (double)c->ihw2);
        const double s_out = std::sqrt(S.sig0old * (double)c->ihw2);
        int red, nob, nun;
        if (c->pass == 0 && c->fl.robust) {
            red = S.redundancynew; nob = S.nrobsnew; nun = S.nrunknownsnew;
        } else {
            red = S.redundancy; nob = S.nrobs; nun = S.nrunknowns;
        }
        std::snprintf(buf, sizeof buf, "Sigma0: %g %g  %d %d %d  IT: %d  ITT: %d\n", s_in, s_out, red, nob, nun, c->pass,
                      S.itt);
        c->logb(b, buf);
        S.stats.sigma_in[c->pass] = s_in;
        S.stats.sigma_out[c->pass] = s_out;
        S.stats.redundancy[c->pass] = red;
        S.stats.nobs[c->pass] = nob;
        S.stats.nunknowns[c->pass] = nun;
        S.stats.passes_done = c->pass + 1;
    }
    // host copies of the stored cameras for the next EncodeCameras
    HIPCHK(hipMemcpyAsync(c->Pst.data(), c->c_Pst.p, sizeof(double) * c->Pst.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(c->Ist.data(), c->c_Ist.p, sizeof(double) * c->Ist.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->fl.robust && c->pass == 0) {
        // a block that stopped on an error returns before RobustDelete
        for (int b = 0; b < c->nb; ++b)
            if (c->B[b].status != RPBA_OK && c->eoff_h[b + 1] > c->eoff_h[b])
                HIPCHK(hipMemsetAsync(c->d_flagst.p + c->eoff_h[b], 0, (size_t)(c->eoff_h[b + 1] - c->eoff_h[b]), c->st));
        const int rc = c->ds.robust_delete(c->d_flagst.p, c->min_keep, c->d_X.p, c->st);
        if (rc) return rc;
        c->structure_dirty = true;
    }
    c->pass++;
    c->stage = (c->pass >= 2) ? 2 : 0;
    return RPBA_OK;
}

// one lockstep LM iteration of every active block (rbacore.cpp:515-581)
static int lm_iteration(rpba_batch *c)
{
    const int nb = c->nb;
    const bool rw = c->pass == 0 && c->fl.robust;
    std::vector<int> ifl(nb);
    std::vector<unsigned char> cf(c->nc);
    for (int b = 0; b < nb; ++b) {
        BlkState &S = c->B[b];
        ifl[b] = S.iflag;
        if (S.active && rw) {
            S.nrunknowns = S.nrunknownsgiven;
            S.nrobs = S.nrobsgiven;
        }
    }
    for (int k = 0; k < c->nc; ++k) {
        const int b = c->camblk_h[k];
        cf[k] = (unsigned char)((c->B[b].active ? 1 : 0) | (k == c->coff[b] ? 2 : 0));
    }
    int rc = 0;
    rc |= c->d_iflag.put(ifl.data(), ifl.size(), c->st);
    rc |= c->d_camflag.put(cf.data(), cf.size(), c->st);
    if (c->pcg) {
        // the PCG cap of every running block (rbacore.cpp:1848-1849); 0 keeps
        // a block that sits out from iterating
        for (int b = 0; b < nb; ++b) {
            std::memset(&c->h_seg[b], 0, sizeof(SegScal));
            c->h_seg[b].msize = c->B[b].active ? std::max(1000, c->B[b].nrioff) : 0;
        }
        rc |= c->d_segS.put(c->h_seg, (size_t)nb, c->st);
    }
    if (rc) return rc;
    const bool replay = c->graph_on && !c->timing;
    if (replay && !c->graph_valid) {
        rc = capture_graph(c);
        if (rc) return rc;
    }
    if (replay) {
        HIPCHK(hipGraphLaunch(c->gexec[0], c->st));
        c->graph_launches++;
    } else {
        launches_pre(c);
    }
    if (c->pcg) {
        rc = run_pcg(c, replay);
        if (rc) return rc;
        if (replay)
            HIPCHK(hipGraphLaunch(c->gexec[2], c->st));
        else
            launches_post(c);
        HIPCHK(hipMemcpyAsync(c->h_seg, c->d_segS.p, sizeof(SegScal) * (size_t)nb, hipMemcpyDeviceToHost, c->st));
    } else if (!replay) {
        launches_post(c);
    }
    HIPCHK(hipMemcpyAsync(c->h_bsum, c->d_bsum.p, sizeof(double) * 3 * (size_t)nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(c->h_nanb, c->d_nanb.p, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->pcg) {
        c->pcg_expect = 0;
        double sb = 0.;
        for (int b = 0; b < nb; ++b)
            if (c->B[b].active) {
                c->pcg_expect = std::max(c->pcg_expect, c->h_seg[b].iters);
                sb += (double)c->h_seg[b].iters * (double)c->sblk_h[b];
            }
        // the SpMV launches of this PCG (one per iteration of its slowest
        // block, exactly so with timing on): each reads the upper blocks of
        // the running blocks' cameras and every diagonal block (648 bytes a
        // block), d (72 per camera), and writes Ad (72) and row partials (8)
        if (c->timing) c->spmv_bytes += 648. * sb + 800. * (double)c->nc * (double)c->pcg_expect;
    }
    if (c->timing) c->resolve_timing();
    std::vector<unsigned char> acc(nb, 0);
    bool any_nan = false;
    std::vector<int> nanb(nb, 0);
    for (int b = 0; b < nb; ++b) {
        BlkState &S = c->B[b];
        if (!S.active) continue;
        if (c->h_nanb[b]) {
            // the reference returns before decoding (rbacore.cpp:541-546)
            nanb[b] = 1;
            any_nan = true;
            c->logb(b, "BA-Error: NANs in result vector\n");
            S.status = RPBA_E_NAN_RESULT;
            S.active = false;
            continue;
        }
        if (rw) {
            S.nrobs -= 2 * (int)c->h_bsum[3 * b + 2];
            S.redundancy = S.nrobs - S.nrunknowns;
            S.wsum = (double)(2 * S.imagepointnr) / c->h_bsum[3 * b + 1];
            S.wsumg = S.wsum;
        }
        S.sig0 = c->h_bsum[3 * b] * S.wsumg / (double)S.redundancy;
        if (S.sig0old < S.sig0 * S.frac) --S.iflag;
        int a = 0;
        if (S.sig0 <= S.sig0old || S.iter == 1) {
            acc[b] = 1;
            S.nrobsnew = S.nrobs;
            S.nrunknownsnew = S.nrunknowns;
            S.redundancynew = S.redundancy;
            S.sig0old = S.sig0;
            a = 1;
        }
        const int pit = c->pcg ? c->h_seg[b].iters : 0;
        S.tr_sig0.push_back(S.sig0);
        S.tr_acc.push_back(a);
        S.tr_nrobs.push_back(S.nrobs);
        S.tr_pcg.push_back(pit);
        S.pcg_total += pit;
        if (S.itt > 200) {  // rbacore.cpp:575-576
            S.active = false;
            continue;
        }
        ++S.itt;
        ++S.iter;
        if (S.iflag <= 0) S.active = false;
    }
    if (any_nan) {
        rc = c->d_nanb.put(nanb.data(), nanb.size(), c->st);
        if (rc) return rc;
        launch_batch_restore_in(c->bargs(), c->d_nanb.p, c->d_In_bak.p, c->c_In.p, c->st);
    }
    rc = c->d_accept.put(acc.data(), acc.size(), c->st);
    if (rc) return rc;
    launch_batch_store(c->bargs(), c->d_accept.p, c->cam(), c->d_Xn.p, c->d_X.p, c->d_flag.p, c->d_flagst.p, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    c->iters_total++;
    bool any = false;
    for (int b = 0; b < nb; ++b) any |= c->B[b].active;
    return any ? 0 : 1;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" void rpba_batch_destroy(rpba_batch *c)
{
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    c->drop_graph();
    c->resolve_timing();
    c->ds.release();
    c->d_wx.release(); c->d_wy.release(); c->d_wxy.release(); c->d_key.release(); c->d_s00.release();
    c->d_flag.release(); c->d_flagst.release(); c->d_camflag.release(); c->d_accept.release();
    c->d_Wrec.release(); c->d_Yrec.release();
    c->d_X.release(); c->d_Xn.release(); c->d_X0.release(); c->d_Vinv.release(); c->d_gp.release();
    c->d_Sd.release(); c->d_b.release(); c->d_sol.release(); c->d_In_bak.release(); c->d_Su.release();
    c->d_Dd.release(); c->d_terms.release(); c->d_bsum.release(); c->d_partial.release();
    c->d_coff.release(); c->d_uptoff.release(); c->d_poff.release(); c->d_eoff.release(); c->d_camblk.release(); c->d_ptblk.release();
    c->d_iflag.release(); c->d_nanb.release(); c->d_ipartial.release(); c->d_iscal.release(); c->d_scal.release();
    c->d_bpart.release(); c->d_npart.release();
    c->c_Rb.release(); c->c_Cb.release(); c->c_Rn.release(); c->c_Cn.release(); c->c_In.release();
    c->c_th.release(); c->c_thn.release(); c->c_Pst.release(); c->c_Ist.release(); c->c_mask.release();
    c->d_pcg.release();
    for (DBuf<double> *v : {&c->d_Cinv, &c->d_px, &c->d_best, &c->d_r, &c->d_dv, &c->d_Ad, &c->d_h, &c->d_h1, &c->d_Sdf,
                            &c->d_ev, &c->d_rowpart, &c->d_colpart, &c->d_dpart})
        v->release();
    c->d_segS.release();
    c->d_ndone.release();
    c->d_live.release();
    if (c->h_bsum) (void)hipHostFree(c->h_bsum);
    if (c->h_nanb) (void)hipHostFree(c->h_nanb);
    if (c->h_ndone) (void)hipHostFree(c->h_ndone);
    if (c->h_seg) (void)hipHostFree(c->h_seg);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

extern "C" int rpba_batch_create(rpba_batch **out, int32_t nblocks, rpba_block *blocks, const rpba_flags *flags,
                                 rpba_batch_log_fn log, void *user)
{
    if (!out || nblocks < 1 || !blocks || !flags) return RPBA_E_INVALID;
    if (flags->precision < 0 || flags->precision > 1 || flags->schur != 0 || flags->solver < 0 || flags->solver > 1 ||
        flags->preconditioner != 0)
        return RPBA_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RPBA_E_HIP;
    rpba_batch *c = new rpba_batch();
    c->nb = nblocks;
    c->io = blocks;
    c->fl = *flags;
    c->pcg = flags->solver == 0;
    c->log = log;
    c->user = user;
    int rc = batch_init(c);
    if (!rc) rc = restore_initial(c);
    if (rc) {
        rpba_batch_destroy(c);
        return rc;
    }
    *out = c;
    return RPBA_OK;
}

extern "C" int rpba_batch_run(rpba_batch *c, int32_t max_iters, int32_t *iters_done, int32_t *finished)
{
    if (!c) return RPBA_E_INVALID;
    int done = 0, status = RPBA_OK;
    while (c->stage != 2 && (max_iters < 0 || done < max_iters)) {
        if (c->stage == 0) {
            status = pass_setup(c);
            if (status) break;
            bool any = false;
            for (auto &S : c->B) any |= S.active;
            if (!any) {
                status = pass_end(c);
                if (status) break;
            }
            continue;
        }
        const int rc = lm_iteration(c);
        ++done;
        if (rc < 0) {
            status = rc;
            break;
        }
        if (rc == 1) {
            status = pass_end(c);
            if (status) break;
        }
    }
    if (status) {
        for (auto &S : c->B)
            if (S.status == RPBA_OK) S.status = status;
        c->stage = 2;
    }
    if (iters_done) *iters_done = done;
    if (finished) *finished = c->stage == 2 ? 1 : 0;
    return status;
}

extern "C" int rpba_batch_reset(rpba_batch *c)
{
    if (!c) return RPBA_E_INVALID;
    HIPCHK(hipStreamSynchronize(c->st));
    const int rc = c->ds.restore_tracks(c->st);
    if (rc) return rc;
    return restore_initial(c);
}

// OutputLSAc per block (rbacore.cpp:644-666): stored R|t and points,
// intrinsics of the last trial, kept observations and points
extern "C" int rpba_batch_download(rpba_batch *c, rpba_stats *stats, int32_t *status)
{
    if (!c) return RPBA_E_INVALID;
    std::vector<int> toff, nrpv, tfo, tcam;
    int rc = c->ds.download_tracks(toff, nrpv, tfo, tcam, c->st);
    if (rc) return rc;
    const int nc = c->nc, npts = c->ds.npts;
    std::vector<double> X(3 * (size_t)npts), P(12 * (size_t)nc), In(3 * (size_t)nc), Is(3 * (size_t)nc);
    if (npts) HIPCHK(hipMemcpyAsync(X.data(), c->d_X.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(P.data(), c->c_Pst.p, sizeof(double) * P.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(In.data(), c->c_In.p, sizeof(double) * In.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(Is.data(), c->c_Ist.p, sizeof(double) * Is.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    std::vector<double> ss(c->nb, 0.);
    std::vector<int64_t> sn(c->nb, 0);
    std::vector<int> kept_pts(c->nb, 0), kept_obs(c->nb, 0);
    for (int b = 0; b < c->nb; ++b) {
        rpba_block &io = c->io[b];
        if (io.obs_kept) std::memset(io.obs_kept, 0, (size_t)io.n_obs);
        if (io.pt_kept) std::memset(io.pt_kept, 0, (size_t)io.n_pts);
    }
    int b = 0;
    for (int i = 0; i < npts; ++i) {
        const int up = nrpv[i];
        while (up >= c->uptoff[b + 1]) ++b;
        rpba_block &io = c->io[b];
        const int p = up - c->uptoff[b];
        for (int j = 0; j < 3; ++j) io.pts[3 * (size_t)p + j] = (float)X[3 * (size_t)i + j];
        if (io.pt_kept) io.pt_kept[p] = 1;
        kept_pts[b]++;
        for (int q = toff[i]; q < toff[(size_t)i + 1]; ++q) {
            const int fo = tfo[q] - c->uobsoff[b];
            if (io.obs_kept) io.obs_kept[fo] = 1;
            kept_obs[b]++;
            // RMS over the kept observations of the stored state
            const int k = tcam[q];
            const double *Pk = P.data() + 12 * (size_t)k;
            const double *x = X.data() + 3 * (size_t)i;
            const double e0 = Pk[0] * x[0] + Pk[1] * x[1] + Pk[2] * x[2] + Pk[3];
            const double e1 = Pk[4] * x[0] + Pk[5] * x[1] + Pk[6] * x[2] + Pk[7];
            const double e2 = Pk[8] * x[0] + Pk[9] * x[1] + Pk[10] * x[2] + Pk[11];
            const double u = e0 / e2, v = e1 / e2, rho = u * u + v * v;
            const double sc = 1. + Is[3 * (size_t)k] * rho + Is[3 * (size_t)k + 1] * rho * rho;
            const double r0 = c->U.obs_xy[2 * (size_t)tfo[q]] - Is[3 * (size_t)k + 2] * sc * u;
            const double r1 = c->U.obs_xy[2 * (size_t)tfo[q] + 1] - Is[3 * (size_t)k + 2] * sc * v;
            ss[b] += r0 * r0 + r1 * r1;
            sn[b] += 1;
        }
    }
    for (int bb = 0; bb < c->nb; ++bb) {
        rpba_block &io = c->io[bb];
        const int k0 = c->coff[bb];
        for (int k = 0; k < io.n_cams; ++k) {
            for (int j = 0; j < 12; ++j) io.cam_P[12 * (size_t)k + j] = (float)P[12 * (size_t)(k0 + k) + j];
            io.cam_d[2 * (size_t)k] = (float)In[3 * (size_t)(k0 + k)];
            io.cam_d[2 * (size_t)k + 1] = (float)In[3 * (size_t)(k0 + k) + 1];
            io.cam_f[k] = (float)In[3 * (size_t)(k0 + k) + 2];
        }
        const BlkState &S = c->B[bb];
        if (status) status[bb] = S.status;
        if (stats) {
            rpba_stats &st = stats[bb];
            rpba_stats keep = st;
            st = S.stats;
            st.itt = S.itt;
            st.pcg_iters = S.pcg_total;
            st.sig0_final = S.sig0old;
            st.rms_px = sn[bb] ? std::sqrt(ss[bb] / (double)(2 * sn[bb])) : 0.;
            st.n_obs_kept = kept_obs[bb];
            st.n_pts_kept = kept_pts[bb];
            st.trace_cap = keep.trace_cap;
            st.trace_sig0 = keep.trace_sig0;
            st.trace_accept = keep.trace_accept;
            st.trace_pcg = keep.trace_pcg;
            st.trace_nrobs = keep.trace_nrobs;
            const int nt = std::min<int>((int)S.tr_sig0.size(), keep.trace_cap);
            st.trace_len = keep.trace_sig0 ? nt : 0;
            for (int j = 0; j < st.trace_len; ++j) {
                st.trace_sig0[j] = S.tr_sig0[j];
                if (st.trace_accept) st.trace_accept[j] = S.tr_acc[j];
                if (st.trace_pcg) st.trace_pcg[j] = S.tr_pcg[j];
                if (st.trace_nrobs) st.trace_nrobs[j] = S.tr_nrobs[j];
            }
        }
    }
    return RPBA_OK;
}

extern "C" int rpba_batch_set_graph(rpba_batch *c, int32_t on)
{
    if (!c) return RPBA_E_INVALID;
    c->graph_on = on != 0;
    c->drop_graph();
    return RPBA_OK;
}

extern "C" int rpba_batch_set_timing(rpba_batch *c, int32_t on)
{
    if (!c) return RPBA_E_INVALID;
    c->timing = on != 0;
    c->ktime.clear();
    c->spmv_bytes = 0.;
    return RPBA_OK;
}

extern "C" int rpba_batch_spmv_bytes(rpba_batch *c, double *bytes)
{
    if (!c || !bytes) return RPBA_E_INVALID;
    *bytes = c->spmv_bytes;
    return RPBA_OK;
}

extern "C" int rpba_batch_kernel_time(rpba_batch *c, const char *kernel, double *total_ms, int64_t *launches)
{
    if (!c || !kernel) return RPBA_E_INVALID;
    auto it = c->ktime.find(kernel);
    if (total_ms) *total_ms = it == c->ktime.end() ? 0. : it->second.first;
    if (launches) *launches = it == c->ktime.end() ? 0 : it->second.second;
    return RPBA_OK;
}

extern "C" int rpba_batch_tuning(rpba_batch *c, char *buf, int32_t cap)
{
    if (!c || !buf || cap < 1) return RPBA_E_INVALID;
    char extra[96];
    if (c->pcg) {
        std::snprintf(buf, (size_t)cap, "segmented block-Jacobi PCG; graph launches %lld", (long long)c->graph_launches);
        return RPBA_OK;
    }
    std::snprintf(extra, sizeof extra, "; graph launches %lld; flops/factorisation %.3g",
                  (long long)c->graph_launches, c->dense.flops());
    std::snprintf(buf, (size_t)cap, "%s%s", c->dense.tuning().c_str(), extra);
    return RPBA_OK;
}

extern "C" int rpba_batch_sizes(rpba_batch *c, int64_t out[6])
{
    if (!c || !out) return RPBA_E_INVALID;
    out[0] = c->nc;
    out[1] = c->ds.npts;
    out[2] = c->ds.ntr;
    out[3] = c->ds.npairs;
    out[4] = c->ds.nblk;
    out[5] = c->dense.matrix_doubles();
    return RPBA_OK;
}

extern "C" int rpba_calc_plsa_batch(int32_t nblocks, rpba_block *blocks, const rpba_flags *flags, rpba_batch_log_fn log,
                                    void *user, rpba_stats *stats, int32_t *status)
{
    rpba_batch *c = nullptr;
    int rc = rpba_batch_create(&c, nblocks, blocks, flags, log, user);
    if (rc) return rc;
    rc = rpba_batch_run(c, -1, nullptr, nullptr);
    const int rd = rpba_batch_download(c, stats, status);
    rpba_batch_destroy(c);
    return rc ? rc : rd;
}
