// Host side of the device solve: the LM state machine of LSA::Adjust
// (rbacore.cpp:333-616), pass set-up (EncodeCameras 999-1087, PrepareWeights2D
// 704-783), RobustDelete (922-995), OutputLSAc (644-666) and the C ABI of
// include/rpba.h.  Every per-observation, per-point and per-camera step runs
// in kernels.hip; the host keeps the scalars of the state machine and reads
// back a few of them per LM iteration.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rpba.h"
#include "dense.h"
#include "host_util.h"
#include "kernels.h"
#include "lp.h"
#include "partition.h"
#include "structure.h"
#include "structure_dev.h"

using namespace rpba;


enum Stage { ST_SETUP = 0, ST_ITER = 1, ST_DONE = 2 };

// pinned host slots for the small device-to-host readbacks (an async copy
// into pageable memory is a blocking staged copy)
struct Readback {
    double sc[2];
    int nflag, nanf;
    PcgScalars pcg[2];  // PCG scalars after chunk n, slot n & 1
    double cost, ws;
};

struct rpba_ctx {
    rpba_block *io = nullptr;
    Readback *hb = nullptr;          // pinned
    hipEvent_t pcg_ev[2] = {nullptr, nullptr};  // end of PCG chunk n, slot n & 1
    hipStream_t st2 = nullptr;                   // off-diagonal Schur blocks, beside st
    hipEvent_t ev_rec = nullptr, ev_pairs = nullptr;
    double *hcam = nullptr;          // pinned, 15 per camera: stored [R|t] (12) and intrinsics (3)
    rpba_flags fl{};
    rpba_log_fn log = nullptr;
    void *user = nullptr;
    hipStream_t st = nullptr;
    int dev = 0;

    // problem
    int nc = 0, nraddpi = 0, addf[3] = {0, 0, 0};
    int normcamera = -1, esvmaxm = 0, lflflag = 0, minimagenrdel = 2;
    int nrioff = 0, nrioff3 = 0;
    float ihw2 = 1.f;
    double defdiff = 5000.;
    int npts_init = 0;               // points kept by PointXXvIn (d_X0 holds them)
    std::vector<double> Pst, Ist;    // stored cameras (host copy at pass boundaries)
    std::vector<double> Pst_init, Ist_init, R0, C0;
    bool structure_dirty = true;     // pairs/blocks must be rebuilt
    int rec32 = 0;                   // FP32 records (precision 1: W/Y; precision 2: whitened Jacobians)
    int jform = 0;                   // whitened Jacobian records (precision 2 and 3; matrix-free)
    bool implicit = false;           // matrix-free S product (schur 1)
    DevStruct ds;                    // device tracks, camera-major order, pairs, blocks
    DenseSolver dense;               // dense Cholesky form (solver 1)

    // device buffers
    DBuf<float> d_wx, d_wy, d_wxy, d_key, d_s00;
    DBuf<unsigned char> d_flag, d_flagst;
    DBuf<double> d_X, d_Xn, d_X0, d_Vinv, d_gp, d_Sd, d_b, d_Cinv, d_In_bak;
    DBuf<unsigned char> d_Wrec, d_Yrec;  // element type by precision
    DBuf<double> d_Su;
    DBuf<double> d_x, d_best, d_r, d_d, d_Ad, d_h, d_h1, d_partial, d_scal, d_Dd, d_ppart, d_rowpart, d_colpart, d_dpart;
    DBuf<double> d_t, d_cpart, d_Dk;  // matrix-free form
    DBuf<unsigned char> d_Jpm, d_Jcm;  // whitened Jacobian records (precision 2)
    DBuf<double> d_dmp;                // damping of S_kk's diagonal (precision 2)
    DBuf<double> d_ptrec;              // per point X, V^-1, g_p as one line (whitened and matrix-free forms)
    // camera-major entry table (point, observation, weights; chunk -> image),
    // rebuilt when the structure (cm_full) or the weights (cm_w) change
    DBuf<int> d_cmpt, d_chcam;
    DBuf<float> d_cmox, d_cmoy, d_cmkey;
    DBuf<float4> d_cmw;
    bool cm_full = true, cm_w = true;
    DBuf<int> d_ipartial, d_iscal;
    DBuf<double> c_Rb, c_Cb, c_Rn, c_Cn, c_In, c_th, c_thn, c_Pst, c_Ist;
    DBuf<unsigned> c_mask;
    DBuf<PcgScalars> d_pcg;
    DBuf<float> d_omega;             // weight of the corrected preconditioner (preconditioner 1, 2)
    // two PCG iterations (launch indices of parities 0 and 1) captured in a
    // hipGraph; the kernels see only the parity of the launch index, so one
    // graph serves every chunk that starts at an even index.  Recaptured when
    // the argument block changes (a new structure, another form).
    hipGraph_t pcg_graph = nullptr;
    hipGraphExec_t pcg_gexec = nullptr;
    PcgArgs pcg_graph_args{};
    bool pcg_graph_on = false;       // measured no faster than launches (rpba_ctx_set_graph)
    float omega_last = 0.f;          // the weight the last PCG ran with (host copy)
    rpba_lp *lp = nullptr;           // graph network of preconditioner 2 (caller-owned)
    // the camera graph of S (features, symmetric CSR) for the network
    DBuf<float> lp_nrm, lp_eb, lp_feat, lp_adj_w;
    DBuf<int> lp_adj_off, lp_adj_nbr, lp_goff;

    // LM state
    int stage = ST_SETUP, it = 0, iflag = 0, iter = 0, itt = 0;
    bool ittflag = true;
    double sig0 = 0, sig0old = 0, sig0in = 0, wsum = 0, wsumg = 0, mult = 1.e-3, frac = 1.05;
    int imagepointnr = 0, nrobs = 0, nrunknowns = 0, redundancy = 0;
    int nrobsnew = 0, nrunknownsnew = 0, redundancynew = 0, nrunknownsgiven = 0, nrobsgiven = 0;
    int status = RPBA_OK;
    int64_t pcg_total = 0;
    int pcg_last = 0;
    int pcg_budget = 0;  // 0: the reference's cap; else every PCG stops there and the run reports it
    rpba_stats stats{};
    std::vector<double> tr_sig0;
    std::vector<int> tr_acc, tr_pcg, tr_nrobs;

    // timing
    bool timing = false;
    std::vector<Timed> pending;
    std::vector<hipEvent_t> evpool;
    std::map<std::string, std::pair<double, int64_t>> ktime;

    CamPtrs cam() const
    {
        CamPtrs c;
        c.Rb = c_Rb.p; c.Cb = c_Cb.p; c.Rn = c_Rn.p; c.Cn = c_Cn.p; c.In = c_In.p;
        c.th = c_th.p; c.thn = c_thn.p; c.Pst = c_Pst.p; c.Ist = c_Ist.p; c.mask = c_mask.p;
        return c;
    }

    void logs(const char *line)
    {
        if (log) log(line, user);
    }

    hipEvent_t ev()
    {
        if (!evpool.empty()) {
            hipEvent_t e = evpool.back();
            evpool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    template <class F>
    void timed(const char *name, F f, int pcg_j = -1)
    {
        if (!timing) {
            f();
            return;
        }
        Timed t{name, ev(), ev(), pcg_j, false};
        (void)hipEventRecord(t.a, st);
        f();
        (void)hipEventRecord(t.b, st);
        pending.push_back(t);
    }
    void resolve_timing()
    {
        for (auto &t : pending) {
            float ms = 0.f;
            if (!t.skip && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
                auto &e = ktime[t.name];
                e.first += ms;
                e.second += 1;
            }
            evpool.push_back(t.a);
            evpool.push_back(t.b);
        }
        pending.clear();
    }
};

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------

static int restore_points(rpba_ctx *c);

// tracks, observation values and original weights of the initial block, and
// the accepted points (LSA construction, rbacore.cpp:37-123)
// RPBA_SETUP_TRACE=1: the set-up and download steps with their wall times on
// stderr (the stream synchronised at each mark)
static void setup_mark(rpba_ctx *c, const char *what)
{
    static const bool on = [] {
        const char *e = std::getenv("RPBA_SETUP_TRACE");
        return e && e[0] == '1';
    }();
    if (!on) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    static auto t0 = std::chrono::steady_clock::now();
    const auto t = std::chrono::steady_clock::now();
    std::fprintf(stderr, "rpba setup: %-28s %9.1f ms\n", what, std::chrono::duration<double, std::milli>(t - t0).count());
    t0 = t;
}

// PointXXvIn on the device from the block's file-order arrays (tracks, the
// observations and original weights gathered per entry, the kept points)
static int upload_tracks(rpba_ctx *c)
{
    if (c->d_X0.alloc(3 * (size_t)std::max(c->io->n_pts, 1))) return RPBA_E_NOMEM;
    int rc = c->ds.build_tracks_device(c->io, c->d_X0.p, c->st);
    if (rc) return rc;
    setup_mark(c, "tracks on the device");
    c->npts_init = c->ds.npts;
    rc |= c->d_X.alloc(3 * (size_t)c->npts_init);
    rc |= c->d_Xn.alloc(3 * (size_t)c->npts_init);
    if (rc) return RPBA_E_NOMEM;
    return restore_points(c);
}

// accepted and trial points <- the initial points (device-to-device)
static int restore_points(rpba_ctx *c)
{
    const size_t n = 3 * (size_t)c->npts_init;
    if (n) {
        HIPCHK(hipMemcpyAsync(c->d_X.p, c->d_X0.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_Xn.p, c->d_X0.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
    }
    HIPCHK(hipStreamSynchronize(c->st));
    c->structure_dirty = true;
    return RPBA_OK;
}

// camera-major order, pairs and blocks on the device, and the per-entry /
// per-block work arrays sized for them.  Returns -1000 - k when image k has
// no observations.
static int build_structure(rpba_ctx *c)
{
    int empty = -1;
    c->ds.spmv_waves = 4 * spmv_grid(0);
    int rc = c->ds.build(c->st, &empty);
    if (rc) return rc;
    if (empty >= 0) return -1000 - empty;
    const size_t ntr = (size_t)c->ds.ntr, np = (size_t)c->ds.npts;
    rc |= c->d_wx.alloc(ntr);
    rc |= c->d_wy.alloc(ntr);
    rc |= c->d_wxy.alloc(ntr);
    rc |= c->d_key.alloc(ntr);
    rc |= c->d_flag.alloc(ntr);
    rc |= c->d_flagst.alloc(ntr);
    const size_t esz = c->rec32 ? sizeof(float) : sizeof(double);
    if (c->jform) {
        rc |= c->d_Jpm.alloc((size_t)kJRecElems * esz * ntr);
        rc |= c->d_Jcm.alloc((size_t)kJRecElems * esz * ntr);
        rc |= c->d_dmp.alloc(9 * (size_t)c->nc);
        rc |= c->d_ptrec.alloc(16 * std::max<size_t>(np, 1));
    } else {
        rc |= c->d_Wrec.alloc((size_t)kRec * esz * ntr);
        rc |= c->d_Yrec.alloc((size_t)kRec * esz * ntr);
        if (c->implicit) rc |= c->d_ptrec.alloc(16 * std::max<size_t>(np, 1));
    }
    rc |= c->d_Vinv.alloc(9 * np);
    rc |= c->d_gp.alloc(3 * np);
    rc |= c->d_Su.alloc(81 * (size_t)std::max(c->ds.nblk, 1));
    if (!c->jform) rc |= c->d_Dd.alloc(90 * (size_t)std::max(c->ds.nch, 1));
    rc |= c->d_rowpart.alloc(63 * (size_t)std::max(c->ds.nchunk, 1));
    rc |= c->d_colpart.alloc(9 * (size_t)std::max(c->ds.nslab, 1));
    rc |= c->d_dpart.alloc((size_t)std::max(spmv_grid(c->ds.nblk), c->ds.nch + kMfDsumGroups));
    rc |= c->d_cmpt.alloc(std::max<size_t>(ntr, 1));
    rc |= c->d_cmox.alloc(std::max<size_t>(ntr, 1));
    rc |= c->d_cmoy.alloc(std::max<size_t>(ntr, 1));
    rc |= c->d_cmw.alloc(std::max<size_t>(ntr, 1));
    rc |= c->d_cmkey.alloc(std::max<size_t>(ntr, 1));
    rc |= c->d_chcam.alloc((size_t)std::max(c->ds.nch, 1));
    c->cm_full = true;
    if (c->implicit) {
        rc |= c->d_t.alloc(3 * std::max<size_t>(np, 1));
        rc |= c->d_cpart.alloc(9 * (size_t)std::max(c->ds.nch, 1));
        if (!c->jform) rc |= c->d_Dk.alloc(81 * (size_t)c->nc);
    }
    const int nb = std::max(cost_blocks(c->ds.ntr), (int)((np + 255) / 256 + 1));
    rc |= c->d_partial.alloc((size_t)nb);
    rc |= c->d_ipartial.alloc((size_t)nb);
    if (rc) return RPBA_E_NOMEM;
    c->structure_dirty = false;
    return RPBA_OK;
}

// EncodeCameras on the host: rectify the stored rotations (k >= 1), base
// R and C = R^T t, zero pose increments, intrinsics into theta, the gauge
// camera's fixed C component (rbacore.cpp:999-1087)
static int encode_cameras(rpba_ctx *c)
{
    const int nc = c->nc;
    std::vector<double> Rb(9 * (size_t)nc), Cb(3 * (size_t)nc), th(9 * (size_t)nc, 0.);
    std::vector<unsigned> mask((size_t)nc, 0);
    for (int k = 0; k < nc; ++k) {
        double *P = c->Pst.data() + 12 * (size_t)k;
        if (k > 0) rectify34(P);
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) Rb[9 * (size_t)k + r * 3 + cc] = P[r * 4 + cc];
        const double *R = Rb.data() + 9 * (size_t)k;
        for (int cc = 0; cc < 3; ++cc)
            Cb[3 * (size_t)k + cc] = R[0 * 3 + cc] * P[3] + R[1 * 3 + cc] * P[7] + R[2 * 3 + cc] * P[11];
        unsigned m = 0;
        if (k > 0) m |= 0x3fu;
        if (c->addf[0]) m |= 1u << 6;
        if (c->addf[1]) m |= 1u << 7;
        if (c->addf[2]) m |= 1u << 8;
        if (k > 0 && k == c->normcamera) {
            const double *C = Cb.data() + 3 * (size_t)k;
            float maxv = 0.f;
            int maxi = 0;
            for (int i = 0; i < 3; ++i)
                if (std::fabs(C[i]) > maxv) {
                    maxv = (float)std::fabs(C[i]);
                    maxi = i;
                }
            c->esvmaxm = maxi;
            m &= ~(1u << (3 + maxi));
        }
        mask[k] = m;
        th[9 * (size_t)k + 6] = c->Ist[3 * (size_t)k + 0];
        th[9 * (size_t)k + 7] = c->Ist[3 * (size_t)k + 1];
        th[9 * (size_t)k + 8] = c->Ist[3 * (size_t)k + 2];
    }
    c->nrioff3 = c->nrioff + 3 * c->ds.npts;
    int rc = 0;
    rc |= c->c_Rb.put(Rb.data(), Rb.size(), c->st);
    rc |= c->c_Cb.put(Cb.data(), Cb.size(), c->st);
    rc |= c->c_th.put(th.data(), th.size(), c->st);
    rc |= c->c_thn.put(th.data(), th.size(), c->st);
    rc |= c->c_mask.put(mask.data(), mask.size(), c->st);
    rc |= c->c_Pst.put(c->Pst.data(), c->Pst.size(), c->st);
    rc |= c->c_Ist.put(c->Ist.data(), c->Ist.size(), c->st);
    if (rc) return RPBA_E_NOMEM;
    HIPCHK(hipStreamSynchronize(c->st));  // host vectors above go out of scope
    return RPBA_OK;
}

static void reset_state(rpba_ctx *c)
{
    c->stage = ST_SETUP;
    c->it = 0;
    c->itt = 0;
    c->iter = 0;
    c->iflag = 0;
    c->ittflag = true;
    c->mult = 1.e-3;
    c->frac = c->lflflag ? (double)1.01f : (double)1.05f;
    c->status = RPBA_OK;
    c->pcg_total = 0;
    c->pcg_last = 0;
    std::memset(&c->stats, 0, sizeof(c->stats));
    c->tr_sig0.clear();
    c->tr_acc.clear();
    c->tr_pcg.clear();
    c->tr_nrobs.clear();
}

// trial pose of image 0 (PrepareIteration rbacore.cpp:151-155): never decoded
static int upload_initial_cameras(rpba_ctx *c)
{
    const int nc = c->nc;
    std::vector<double> Rn(9 * (size_t)nc, 0.), Cn(3 * (size_t)nc, 0.), In(3 * (size_t)nc);
    for (int j = 0; j < 9; ++j) Rn[j] = c->R0[j];
    for (int j = 0; j < 3; ++j) Cn[j] = c->C0[j];
    for (size_t j = 0; j < In.size(); ++j) In[j] = c->Ist[j];
    int rc = 0;
    rc |= c->c_Rn.put(Rn.data(), Rn.size(), c->st);
    rc |= c->c_Cn.put(Cn.data(), Cn.size(), c->st);
    rc |= c->c_In.put(In.data(), In.size(), c->st);
    // the stored state (read by rpba_ctx_download) exists on the device from
    // here on, also for a solve that stops before its first pass set-up
    rc |= c->c_Pst.put(c->Pst.data(), c->Pst.size(), c->st);
    rc |= c->c_Ist.put(c->Ist.data(), c->Ist.size(), c->st);
    if (rc) return RPBA_E_NOMEM;
    HIPCHK(hipStreamSynchronize(c->st));
    return RPBA_OK;
}

// ---------------------------------------------------------------------------
// device steps
// ---------------------------------------------------------------------------

static LinArgs lin_args(rpba_ctx *c)
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
    a.Dk = c->implicit ? c->d_Dk.p : nullptr;
    a.yrank = c->implicit ? c->ds.camrank : nullptr;
    // the camera-major entry table (bundle_step refreshes it before the
    // camera kernels run when the structure or the weights changed)
    a.cm_pt = c->d_cmpt.p; a.cm_ox = c->d_cmox.p; a.cm_oy = c->d_cmoy.p;
    a.cm_w = c->d_cmw.p;
    a.jform = c->jform; a.Jpm = c->d_Jpm.p; a.Jcm = c->d_Jcm.p; a.dmp = c->d_dmp.p; a.ptrec = c->d_ptrec.p;
    // matrix-free W / Y records: Y is formed on the camera side (k_lin_chunks)
    a.ycm = c->implicit && !c->jform;
    a.ch4 = c->ds.ch4;
    return a;
}

static PcgArgs pcg_args(rpba_ctx *c)
{
    PcgArgs a;
    std::memset(&a, 0, sizeof a);  // padding too: the PCG graph compares argument blocks bytewise
    a.nc = c->nc;
    a.msize = std::max(1000, c->nrioff);  // rbacore.cpp:1848-1849
    if (c->pcg_budget > 0) a.msize = std::min(a.msize, c->pcg_budget);
    a.Sd = c->d_Sd.p; a.Su = c->d_Su.p; a.Cinv = c->d_Cinv.p; a.b = c->d_b.p;
    a.row_off = c->ds.roff; a.col_off = c->ds.coff; a.col_blk = c->ds.cblk; a.blk_k = c->ds.bk; a.blk_l = c->ds.bl;
    a.x = c->d_x.p; a.best = c->d_best.p; a.r = c->d_r.p; a.d = c->d_d.p; a.Ad = c->d_Ad.p;
    a.h = c->d_h.p; a.h1 = c->d_h1.p; a.part = c->d_ppart.p; a.S = c->d_pcg.p;
    a.sm = c->ds.sumap();
    a.nslab = c->ds.nslab; a.nchunk = c->ds.nchunk; a.nstrip = c->ds.nstrip; a.slab_chunk = c->ds.slab_chunk;
    a.slab_bs = c->ds.slab_bs; a.slab_l = c->ds.slab_l; a.slab_mask = c->ds.slab_mask; a.strip_s0 = c->ds.strip_s0;
    a.chunk_off = c->ds.chunk_off; a.chunk_strip = c->ds.chunk_strip; a.cs_off = c->ds.cs_off; a.cs_list = c->ds.cs_list;
    a.rowpart = c->d_rowpart.p; a.colpart = c->d_colpart.p;
    a.dpart = c->d_dpart.p; a.nblk = c->ds.nblk;
    a.ndpart = c->implicit ? c->ds.nch : spmv_grid(c->ds.nblk);
    a.dsum = c->implicit ? c->d_dpart.p + c->ds.nch : c->d_dpart.p;
    a.ndsum = c->implicit ? kMfDsumGroups : a.ndpart;
    a.j = 0;
    a.implicit = c->implicit ? 1 : 0; a.npts = c->ds.npts; a.ntr = c->ds.ntr; a.nch = c->ds.nch; a.rec32 = c->rec32;
    a.toff = c->ds.toff; a.tcam = c->ds.tcam; a.tpt = c->ds.tpt; a.camq = c->ds.camq; a.camrank = c->ds.camrank;
    a.pgrp = c->ds.pgrp; a.npgrp = c->ds.npgrp;
    a.ch_start = c->ds.ch_start; a.ch_off = c->ds.ch_off;
    a.Wrec = c->d_Wrec.p; a.Yrec = c->d_Yrec.p; a.Dk = c->d_Dk.p; a.tc = c->d_t.p; a.cpart = c->d_cpart.p;
    a.cq_pt = c->implicit ? c->d_cmpt.p : nullptr;
    a.ch_cam = c->implicit ? c->d_chcam.p : nullptr;
    a.omega = c->fl.preconditioner >= 1 ? c->d_omega.p : nullptr;
    a.jform = c->jform; a.Jpm = c->d_Jpm.p; a.Jcm = c->d_Jcm.p; a.dmp = c->d_dmp.p; a.Vinv = c->d_Vinv.p;
    a.pg4 = c->ds.pg4; a.ch4 = c->ds.ch4;
    return a;
}

static StepArgs step_args(rpba_ctx *c, const double *sol)
{
    StepArgs a;
    a.nc = c->nc; a.npts = c->ds.npts; a.cam = c->cam();
    a.toff = c->ds.toff; a.tcam = c->ds.tcam; a.pgrp = c->ds.pgrp; a.npgrp = c->ds.npgrp; a.pg4 = c->ds.pg4;
    a.Wrec = c->d_Wrec.p; a.rec32 = c->rec32;
    a.Vinv = c->d_Vinv.p; a.gp = c->d_gp.p; a.sol = sol; a.X = c->d_X.p; a.Xn = c->d_Xn.p;
    a.S = c->d_pcg.p;
    a.jform = c->jform; a.Jpm = c->d_Jpm.p; a.pg4 = c->ds.pg4;
    a.wfac = c->implicit && !c->jform;  // the matrix-free form keeps W as its factors
    return a;
}

static CostArgs cost_args(rpba_ctx *c)
{
    CostArgs a;
    a.ntr = c->ds.ntr; a.robust = c->fl.robust != 0; a.write_key = 1; a.defdiff = c->defdiff;
    a.cam = c->cam(); a.tcam = c->ds.tcam; a.tpt = c->ds.tpt;
    a.ox = c->ds.ox; a.oy = c->ds.oy; a.wx = c->d_wx.p; a.wy = c->d_wy.p; a.wxy = c->d_wxy.p;
    a.w0x = c->ds.w0x; a.w0y = c->ds.w0y; a.w0xy = c->ds.w0xy;
    a.X = c->d_Xn.p; a.key = c->d_key.p; a.partial = c->d_partial.p;
    // the reduced keys in camera-major order feed the robust reweighting of
    // pass 0 only
    a.cm_key = (c->it == 0 && c->fl.robust) ? c->d_cmkey.p : nullptr;
    a.camrank = c->ds.camrank; a.toff = c->ds.toff;
    return a;
}

// trial cost at (trial cameras, Xn) into d_scal[0]; sets the float keys
// (srmvflag = true)
static void launch_trial_cost(rpba_ctx *c)
{
    CostArgs a = cost_args(c);
    c->timed("cost", [&] { launch_cost(a, c->d_scal.p, c->st); });
}

static int trial_cost(rpba_ctx *c, double &cost)
{
    launch_trial_cost(c);
    HIPCHK(hipMemcpyAsync(&c->hb->cost, c->d_scal.p, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    cost = c->hb->cost;
    return RPBA_OK;
}

static int finish_bundle_step(rpba_ctx *c, const PcgScalars &hs);
static int pcg_solve(rpba_ctx *c, PcgArgs &pa, bool wait_pairs, PcgScalars &hs);
static int lp_predict(rpba_ctx *c);

// CalcBundleSep (rbacore.cpp:1094-1195) + theta_new = theta - x + decode.
// The NaN / inf check of the step (rbacore.cpp:541-546) is raised on the
// device by k_backsub (points) and k_decode (the float camera step); the
// caller reads it with the trial cost and, when it is set, restores the trial
// intrinsics saved here, since the reference returns before decoding.
static int bundle_step(rpba_ctx *c)
{
    LinArgs la = lin_args(c);
    SchurArgs sa;
    sa.nblk = c->ds.nblk; sa.poff = c->ds.poff; sa.pab = c->ds.pab;
    sa.Yrec = c->d_Yrec.p; sa.Wrec = c->d_Wrec.p; sa.Su = c->d_Su.p; sa.sm = c->ds.sumap(); sa.sperm = c->ds.sperm; sa.nslab = c->ds.nslab; sa.slab_bs = c->ds.slab_bs; sa.npairs = c->ds.npairs; sa.rec32 = c->rec32;
    if (c->cm_full || c->cm_w) {
        CmTable ct{c->ds.camrank, c->d_cmpt.p, c->d_cmox.p, c->d_cmoy.p, c->d_cmw.p, c->d_chcam.p};
        c->timed("cm_table", [&] { launch_cm_table(la, ct, c->cm_full, c->st); });
        c->cm_full = c->cm_w = false;
    }
    if (c->jform) {
        // whitened Jacobian records: V, g_p and the point-major records, then
        // the camera blocks with their Schur terms and the camera-major records
        c->timed("lin_points", [&] { launch_lin_points_j(la, c->st); });
        c->timed("lin_cameras", [&] { launch_lin_cameras_j(la, c->st); });
        c->timed("bj_inverse", [&] { launch_bj_inverse(c->nc, c->d_Sd.p, c->d_Cinv.p, c->st); });
        PcgArgs pa = pcg_args(c);
        PcgScalars hs{};
        const int rc = pcg_solve(c, pa, false, hs);
        if (rc) return rc;
        if (c->pcg_budget > 0 && hs.iters >= pa.msize && pa.msize < std::max(1000, c->nrioff)) return RPBA_E_BUDGET;
        return finish_bundle_step(c, hs);
    }
    c->timed("lin_points", [&] { launch_lin_points(la, c->st); });
    // the off-diagonal blocks depend only on the records: without kernel
    // timing they run on a second stream, beside the diagonal terms, the
    // camera blocks, their inverses and the PCG set-up
    const bool side = !c->implicit && !c->timing;
    if (side) {
        HIPCHK(hipEventRecord(c->ev_rec, c->st));
        HIPCHK(hipStreamWaitEvent(c->st2, c->ev_rec, 0));
        launch_schur_pairs(sa, c->st2);
        HIPCHK(hipEventRecord(c->ev_pairs, c->st2));
    } else if (!c->implicit) {
        c->timed("schur_pairs", [&] { launch_schur_pairs(sa, c->st); });
    }
    c->timed("schur_diag", [&] { launch_schur_diag(la, c->st); });
    c->timed("lin_cameras", [&] { launch_lin_cameras(la, c->st); });
    if (c->fl.solver == 1) {
        // exact solve of the same S and b on the matrix cores
        if (side) HIPCHK(hipStreamWaitEvent(c->st, c->ev_pairs, 0));
        c->timed("dense_assemble", [&] {
            c->dense.assemble(c->d_Sd.p, c->d_Su.p, c->ds.sumap(), c->ds.bk, c->ds.bl, c->ds.nblk, c->d_b.p, c->st);
        });
        c->dense.factor_solve(c->d_best.p, c->st, [&](const char *nm, auto f) { c->timed(nm, f); });
        PcgScalars hs{};
        hs.done = 1;
        return finish_bundle_step(c, hs);
    }
    c->timed("bj_inverse", [&] { launch_bj_inverse(c->nc, c->d_Sd.p, c->d_Cinv.p, c->st); });
    PcgArgs pa = pcg_args(c);
    if (pa.omega) {
        // the corrected preconditioner multiplies by the off-diagonal blocks,
        // and its weight is the fixed one or the graph network's prediction
        // from this S
        if (side) HIPCHK(hipStreamWaitEvent(c->st, c->ev_pairs, 0));
        if (c->fl.preconditioner == 2) {
            const int rcl = lp_predict(c);
            if (rcl) return rcl;
        } else {
            HIPCHK(hipMemcpyAsync(c->d_omega.p, &c->fl.pc_omega, sizeof(float), hipMemcpyHostToDevice, c->st));
        }
    }
    PcgScalars hs{};
    const int rc = pcg_solve(c, pa, side && !pa.omega, hs);
    if (rc) return rc;
    if (c->pcg_budget > 0 && hs.iters >= pa.msize && pa.msize < std::max(1000, c->nrioff)) return RPBA_E_BUDGET;
    return finish_bundle_step(c, hs);
}

// SparsePCG (rbacore.cpp:1807-1914) on the device: set-up, then three launches
// per iteration (five with the corrected preconditioner).  wait_pairs: the
// off-diagonal blocks are still being formed on the second stream.
static int pcg_solve(rpba_ctx *c, PcgArgs &pa, bool wait_pairs, PcgScalars &hs)
{
    c->timed("pcg_init", [&] {
        launch_pcg_init(pa, c->st);
        if (pa.omega) launch_pcg_init_w(pa, c->st);
    });
    if (wait_pairs) HIPCHK(hipStreamWaitEvent(c->st, c->ev_pairs, 0));
    // launch j reads the scalars of slot j & 1 and writes slot (j + 1) & 1;
    // once the stopping rule fired, later launches only pass the state on.
    // Chunks are pipelined: chunk n + 1 is queued before the host reads the
    // scalars after chunk n, so the GPU never waits for the host; at most one
    // short chunk runs after the stop.
    hs = PcgScalars{};
    int launched = 0, n = 0;
    const size_t first_timed = c->pending.size();
    const bool use_graph = c->pcg_graph_on && !c->timing;
    if (use_graph) {
        PcgArgs key = pa;
        key.j = 0;
        if (!c->pcg_gexec || std::memcmp(&key, &c->pcg_graph_args, sizeof key) != 0) {
            if (c->pcg_gexec) (void)hipGraphExecDestroy(c->pcg_gexec);
            if (c->pcg_graph) (void)hipGraphDestroy(c->pcg_graph);
            c->pcg_gexec = nullptr;
            c->pcg_graph = nullptr;
            HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
            for (int j = 0; j < 2; ++j) {
                key.j = j;
                launch_spmv(key, c->st);
                if (key.omega)
                    launch_pcg_update_w(key, c->st);
                else
                    launch_pcg_update(key, c->st);
            }
            HIPCHK(hipStreamEndCapture(c->st, &c->pcg_graph));
            HIPCHK(hipGraphInstantiate(&c->pcg_gexec, c->pcg_graph, nullptr, nullptr, 0));
            key.j = 0;
            c->pcg_graph_args = key;
        }
    }
    auto queue_chunk = [&](int len) -> int {
        if (use_graph) {
            len += len & 1;  // whole graph replays: chunks of even length from an even index
            for (int j = 0; j < len; j += 2) HIPCHK(hipGraphLaunch(c->pcg_gexec, c->st));
            launched += len;
            PcgScalars *slot = &c->hb->pcg[n & 1];
            HIPCHK(hipMemcpyAsync(slot, c->d_pcg.p + (launched & 1), sizeof(PcgScalars), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipEventRecord(c->pcg_ev[n & 1], c->st));
            ++n;
            return RPBA_OK;
        }
        for (int j = 0; j < len; ++j) {
            pa.j = launched + j;
            c->timed("spmv", [&] { launch_spmv(pa, c->st); }, pa.j);
            c->timed("pcg_update", [&] {
                if (pa.omega)
                    launch_pcg_update_w(pa, c->st);
                else
                    launch_pcg_update(pa, c->st);
            }, pa.j);
        }
        launched += len;
        PcgScalars *slot = &c->hb->pcg[n & 1];
        HIPCHK(hipMemcpyAsync(slot, c->d_pcg.p + (launched & 1), sizeof(PcgScalars), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipEventRecord(c->pcg_ev[n & 1], c->st));
        ++n;
        return RPBA_OK;
    };
    // an iteration that streams more than ~256 MB (S d) outlasts the host's
    // turnaround by far: chunks of one iteration (after a first chunk of
    // two) stop within a launch of the rule; smaller ones take the previous
    // solve's count first, then chunks of three
    const double it_bytes = c->implicit ? (c->jform ? (c->rec32 ? 200. : 392.) : 450.) * (double)c->ds.ntr : 648. * ((double)c->ds.nblk + (double)c->nc);
    const bool long_it = it_bytes > 2.56e8;
    int rc = queue_chunk(long_it ? 2 : std::max(2, c->pcg_last));
    if (rc) return rc;
    for (int m = 0;; ++m) {
        rc = queue_chunk(use_graph ? 2 : (long_it ? 1 : 3));
        if (rc) return rc;
        HIPCHK(hipEventSynchronize(c->pcg_ev[m & 1]));
        hs = c->hb->pcg[m & 1];
        if (hs.done) break;
        if (launched > pa.msize + 16) return RPBA_E_HIP;  // cannot happen: the kernel stops at msize
    }
    // kernel timing counts the launches that ran an iteration (j < iters)
    for (size_t q = first_timed; q < c->pending.size(); ++q)
        if (c->pending[q].pcg_j >= hs.iters) c->pending[q].skip = true;
    if (pa.omega) HIPCHK(hipMemcpy(&c->omega_last, pa.omega, sizeof(float), hipMemcpyDeviceToHost));
    return RPBA_OK;
}

// the camera graph of the current S into the context's buffers
static int lp_graph(rpba_ctx *c, LpGraphs &G)
{
    const int nc = c->nc, nb = c->ds.nblk;
    int rc = 0;
    rc |= c->lp_nrm.alloc((size_t)nc);
    rc |= c->lp_eb.alloc((size_t)std::max(nb, 1));
    rc |= c->lp_feat.alloc((size_t)nc * kLpF);
    rc |= c->lp_adj_off.alloc((size_t)nc + 1);
    rc |= c->lp_adj_nbr.alloc(2 * (size_t)std::max(nb, 1));
    rc |= c->lp_adj_w.alloc(2 * (size_t)std::max(nb, 1));
    if (!c->lp_goff.p) {
        const int go[2] = {0, nc};
        rc |= c->lp_goff.put(go, 2, c->st);
        HIPCHK(hipStreamSynchronize(c->st));
    }
    if (rc) return RPBA_E_NOMEM;
    lp_graph_from_S(nc, nb, c->d_Sd.p, c->d_Su.p, c->ds.sumap(), c->ds.bk, c->ds.bl, c->ds.roff, c->ds.coff, c->ds.cblk, c->lp_nrm.p,
                    c->lp_eb.p, c->lp_feat.p, c->lp_adj_off.p, c->lp_adj_nbr.p, c->lp_adj_w.p, c->st);
    G.ng = 1;
    G.nn = nc;
    G.ne = 2LL * nb;
    G.goff = c->lp_goff.p;
    G.feat = c->lp_feat.p;
    G.adj_off = c->lp_adj_off.p;
    G.adj_nbr = c->lp_adj_nbr.p;
    G.adj_w = c->lp_adj_w.p;
    G.label = nullptr;
    return RPBA_OK;
}

// preconditioner 2: the network's weight for this S into d_omega (0, i.e.
// block-Jacobi, when no network is attached)
static int lp_predict(rpba_ctx *c)
{
    if (!c->lp) {
        const float z = 0.f;
        HIPCHK(hipMemcpyAsync(c->d_omega.p, &z, sizeof(float), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        return RPBA_OK;
    }
    LpGraphs G;
    int rc = 0;
    c->timed("lp_graph", [&] { rc = lp_graph(c, G); });
    if (rc) return rc;
    c->timed("lp_forward", [&] { rc = lp_model_forward(c->lp, G, c->d_omega.p, c->st); });
    return rc;
}

// after the PCG: back-substitution and decode of the trial state
static int finish_bundle_step(rpba_ctx *c, const PcgScalars &hs)
{
    c->pcg_last = hs.iters;
    c->pcg_total += hs.iters;
    StepArgs sa2 = step_args(c, c->d_best.p);
    HIPCHK(hipMemsetAsync(&c->d_pcg.p->nan_flag, 0, sizeof(int), c->st));
    c->timed("backsub", [&] { launch_backsub(sa2, c->st); });
    HIPCHK(hipMemcpyAsync(c->d_In_bak.p, c->c_In.p, sizeof(double) * 3 * (size_t)c->nc, hipMemcpyDeviceToDevice, c->st));
    c->timed("decode", [&] { launch_decode(sa2, c->st); });
    return RPBA_OK;
}

static int device_store(rpba_ctx *c)
{
    launch_store_cams(c->cam(), c->nc, c->st);
    HIPCHK(hipMemcpyAsync(c->d_X.p, c->d_Xn.p, sizeof(double) * 3 * (size_t)c->ds.npts, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_flagst.p, c->d_flag.p, (size_t)c->ds.ntr, hipMemcpyDeviceToDevice, c->st));
    c->nrobsnew = c->nrobs;
    c->nrunknownsnew = c->nrunknowns;
    c->redundancynew = c->redundancy;
    return RPBA_OK;
}

// ---------------------------------------------------------------------------
// the state machine
// ---------------------------------------------------------------------------

static int pass_setup(rpba_ctx *c)
{
    char buf[256];
    c->sig0old = (double)1.e20f;
    if (c->it > 0) c->mult = 1.e-6;
    c->imagepointnr = c->ds.ntr;
    c->nrobs = 2 * c->imagepointnr;
    if (c->structure_dirty) {
        const int rc = build_structure(c);
        if (rc <= -1000) {
            if (c->nc > 3) {
                std::snprintf(buf, sizeof buf, "BA-Error: No observations in image %d (image index)\n", -1000 - rc);
                c->logs(buf);
            }
            return RPBA_E_NO_OBS_IN_IMAGE;
        }
        if (rc) return rc;
    }
    const size_t ntr = (size_t)c->ds.ntr;
    if (c->ittflag) {
        // PrepareWeights2D: Wij = original weights; wsum over the entries
        HIPCHK(hipMemcpyAsync(c->d_wx.p, c->ds.w0x, sizeof(float) * ntr, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_wy.p, c->ds.w0y, sizeof(float) * ntr, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->d_wxy.p, c->ds.w0xy, sizeof(float) * ntr, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemsetAsync(c->d_flag.p, 0, ntr, c->st));
        HIPCHK(hipMemsetAsync(c->d_flagst.p, 0, ntr, c->st));
        launch_wsum(c->ds.ntr, c->ds.w0x, c->ds.w0y, c->d_partial.p, c->d_scal.p, c->st);
        HIPCHK(hipMemcpyAsync(&c->hb->ws, c->d_scal.p, sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->wsum = c->hb->ws;
        c->ittflag = false;
        c->cm_w = true;
    }
    c->wsum = (double)(2 * c->imagepointnr) / c->wsum;
    c->wsumg = c->wsum;
    int rc = encode_cameras(c);
    if (rc) return rc;
    c->nrunknowns = c->nrioff3 + c->nraddpi * c->nc;  // counts the intrinsics twice (rbacore.cpp:462)
    c->redundancy = c->nrobs - c->nrunknowns;
    if (c->redundancy < 0) {
        if (c->nc > 3) c->logs("BA-Error: Low redundancy\n");
        return RPBA_E_LOW_REDUNDANCY;
    }
    if (c->it == 0) {
        StepArgs sa = step_args(c, nullptr);
        launch_decode(sa, c->st);
        HIPCHK(hipMemcpyAsync(c->d_Xn.p, c->d_X.p, sizeof(double) * 3 * (size_t)c->ds.npts, hipMemcpyDeviceToDevice, c->st));
        double cost = 0.;
        rc = trial_cost(c, cost);
        if (rc) return rc;
        c->sig0 = cost * c->wsumg / (double)c->redundancy;
        c->sig0in = c->sig0old = c->sig0;
    } else {
        c->sig0in = c->sig0old;
    }
    c->iflag = c->fl.robust ? 2 : 1;
    c->iter = 1;
    c->nrunknownsgiven = c->nrunknowns;
    c->nrobsgiven = c->nrobs;
    c->stage = ST_ITER;
    return RPBA_OK;
}

static int pass_end(rpba_ctx *c)
{
    char buf[256];
    const double s_in = std::sqrt(c->sig0in * (double)c->ihw2);
    const double s_out = std::sqrt(c->sig0old * (double)c->ihw2);
    int red, nob, nun;
    if (c->it == 0 && c->fl.robust) {
        red = c->redundancynew; nob = c->nrobsnew; nun = c->nrunknownsnew;
    } else {
        red = c->redundancy; nob = c->nrobs; nun = c->nrunknowns;
    }
    std::snprintf(buf, sizeof buf, "Sigma0: %g %g  %d %d %d  IT: %d  ITT: %d\n", s_in, s_out, red, nob, nun, c->it, c->itt);
    c->logs(buf);
    c->stats.sigma_in[c->it] = s_in;
    c->stats.sigma_out[c->it] = s_out;
    c->stats.redundancy[c->it] = red;
    c->stats.nobs[c->it] = nob;
    c->stats.nunknowns[c->it] = nun;
    c->stats.passes_done = c->it + 1;
    // host copies of the stored cameras for the next EncodeCameras
    double *hP = c->hcam, *hI = c->hcam + c->Pst.size();
    HIPCHK(hipMemcpyAsync(hP, c->c_Pst.p, sizeof(double) * c->Pst.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(hI, c->c_Ist.p, sizeof(double) * c->Ist.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    std::copy(hP, hP + c->Pst.size(), c->Pst.begin());
    std::copy(hI, hI + c->Ist.size(), c->Ist.begin());
    if (c->fl.robust && c->it == 0) {
        const int rc = c->ds.robust_delete(c->d_flagst.p, c->minimagenrdel, c->d_X.p, c->st);
        if (rc) return rc;
        c->structure_dirty = true;
    }
    c->ittflag = true;
    c->it++;
    c->stage = (c->it >= 2) ? ST_DONE : ST_SETUP;
    return RPBA_OK;
}

static int lm_iteration(rpba_ctx *c)
{
    const bool rw = c->it == 0 && c->fl.robust;
    if (rw) {
        c->nrunknowns = c->nrunknownsgiven;
        c->nrobs = c->nrobsgiven;
        RobustArgs ra;
        ra.nc = c->nc; ra.ntr = c->ds.ntr; ra.iflag = c->iflag;
        ra.camoff = c->ds.camoff; ra.camq = c->ds.camq; ra.tpt = c->ds.tpt; ra.toff = c->ds.toff; ra.tcam = c->ds.tcam;
        ra.key = c->d_key.p; ra.w0x = c->ds.w0x; ra.w0y = c->ds.w0y; ra.w0xy = c->ds.w0xy;
        ra.s00 = c->d_s00.p; ra.wx = c->d_wx.p; ra.wy = c->d_wy.p; ra.wxy = c->d_wxy.p;
        ra.flag = c->d_flag.p; ra.partial = c->d_partial.p; ra.ipartial = c->d_ipartial.p;
        // the median reads the keys the last cost wrote in camera-major
        // order; the new weights go to the camera-major table as well
        ra.cm_key = c->d_cmkey.p; ra.cm_w = c->d_cmw.p; ra.camrank = c->ds.camrank;
        c->timed("reweight", [&] { launch_reweight(ra, c->d_scal.p + 1, c->d_iscal.p, c->st); });
    }
    int rc = bundle_step(c);
    if (rc) return rc;
    launch_trial_cost(c);
    // one readback per LM iteration: cost, reweighting sums, NaN flag
    Readback *hb = c->hb;
    HIPCHK(hipMemcpyAsync(hb->sc, c->d_scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(&hb->nflag, c->d_iscal.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(&hb->nanf, &c->d_pcg.p->nan_flag, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    const double sc[2] = {hb->sc[0], hb->sc[1]};
    const int nflag = hb->nflag, nanf = hb->nanf;
    if (nanf) {
        HIPCHK(hipMemcpyAsync(c->c_In.p, c->d_In_bak.p, sizeof(double) * 3 * (size_t)c->nc, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->logs("BA-Error: NANs in result vector\n");
        return RPBA_E_NAN_RESULT;
    }
    if (rw) {
        c->nrobs -= 2 * nflag;
        c->redundancy = c->nrobs - c->nrunknowns;
        c->wsum = (double)(2 * c->imagepointnr) / sc[1];
        c->wsumg = c->wsum;
    }
    const double cost = sc[0];
    c->sig0 = cost * c->wsumg / (double)c->redundancy;
    if (c->sig0old < c->sig0 * c->frac) --c->iflag;
    int acc = 0;
    if (c->sig0 <= c->sig0old || c->iter == 1) {
        rc = device_store(c);
        if (rc) return rc;
        c->sig0old = c->sig0;
        acc = 1;
    }
    c->tr_sig0.push_back(c->sig0);
    c->tr_acc.push_back(acc);
    c->tr_pcg.push_back(c->pcg_last);
    c->tr_nrobs.push_back(c->nrobs);
    if (c->timing) {
        HIPCHK(hipStreamSynchronize(c->st));
        c->resolve_timing();
    }
    if (c->itt > 200) return 1;  // break out of the pass (rbacore.cpp:575-576)
    ++c->itt;
    ++c->iter;
    return (c->iflag > 0) ? 0 : 1;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

static int ctx_init(rpba_ctx *c)
{
    rpba_block *io = c->io;
    const int nc = io->n_cams;
    c->nc = nc;
    // precision 1: FP32 Jacobians and W/Y records, FP64 sums, S, PCG and cost;
    // precision 2: whitened Jacobians rounded to FP32 once, everything else
    // formed from them in FP64, matrix-free product; precision 3: the same
    // records in FP64
    c->rec32 = c->fl.precision == 1 || c->fl.precision == 2;
    c->jform = c->fl.precision >= 2;
    c->implicit = c->fl.schur == 1 || c->jform;
    c->ds.with_pairs = !c->implicit;
    c->addf[0] = c->fl.est_k2 != 0;
    c->addf[1] = c->fl.est_k4 != 0;
    c->addf[2] = c->fl.est_f != 0;
    c->nraddpi = c->addf[0] + c->addf[1] + c->addf[2];
    float ihw2 = 0.f;
    for (int i = 0; i < nc; ++i) {
        const int w = io->img_w ? io->img_w[i] : 1, h = io->img_h ? io->img_h[i] : 1;
        ihw2 += (float)((h + w) * (h + w)) * 0.25f;
    }
    c->ihw2 = ihw2 / (float)nc;
    c->defdiff = 5000. / (double)sqrtf(c->ihw2);  // rbacore.cpp:48
    c->minimagenrdel = nc > 2 ? 3 : 2;
    int rc = 0;
    c->Pst.resize(12 * (size_t)nc);
    c->Ist.resize(3 * (size_t)nc);
    for (int k = 0; k < nc; ++k) {
        for (int j = 0; j < 12; ++j) c->Pst[12 * (size_t)k + j] = (double)io->cam_P[12 * (size_t)k + j];
        c->Ist[3 * (size_t)k] = (double)io->cam_d[2 * (size_t)k];
        c->Ist[3 * (size_t)k + 1] = (double)io->cam_d[2 * (size_t)k + 1];
        c->Ist[3 * (size_t)k + 2] = (double)io->cam_f[k];
    }
    // PrepareCameras: gauge camera from the unrectified matrices (rbacore.cpp:670-700)
    {
        const double *P0 = c->Pst.data();
        double T0[3];
        for (int r = 0; r < 3; ++r) T0[r] = P0[r * 4] * P0[3] + P0[r * 4 + 1] * P0[7] + P0[r * 4 + 2] * P0[11];
        double nd = 0.;
        c->normcamera = nc > 1 ? 1 : -1;
        for (int i = 1; i < nc; ++i) {
            const double *P = c->Pst.data() + 12 * (size_t)i;
            double d = 0.;
            for (int r = 0; r < 3; ++r) {
                const double T = P[r * 4] * P[3] + P[r * 4 + 1] * P[7] + P[r * 4 + 2] * P[11];
                d += (T - T0[r]) * (T - T0[r]);
            }
            if (d > nd) {
                nd = d;
                c->normcamera = i;
            }
        }
    }
    // PrepareIteration: rectify image 0, its trial pose, lflflag, nrioff
    rectify34(c->Pst.data());
    c->R0.assign(9, 0.);
    c->C0.assign(3, 0.);
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) c->R0[r * 3 + cc] = c->Pst[r * 4 + cc];
    for (int cc = 0; cc < 3; ++cc)
        c->C0[cc] = c->R0[0 * 3 + cc] * c->Pst[3] + c->R0[1 * 3 + cc] * c->Pst[7] + c->R0[2 * 3 + cc] * c->Pst[11];
    c->lflflag = 0;
    for (int k = 0; k < nc; ++k)
        if (io->cam_f[k] > 2.f) c->lflflag = 1;
    c->nrioff = c->nraddpi + (5 + c->nraddpi) + (6 + c->nraddpi) * (nc - 2);
    c->Pst_init = c->Pst;
    c->Ist_init = c->Ist;

    setup_mark(c, "host set-up");
    HIPCHK(hipGetDevice(&c->dev));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_rec, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_pairs, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->pcg_ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->pcg_ev[1], hipEventDisableTiming));
    if (hipHostMalloc((void **)&c->hb, sizeof(Readback), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->hcam, sizeof(double) * 15 * (size_t)nc, hipHostMallocDefault) != hipSuccess)
        return RPBA_E_NOMEM;
    const size_t n9 = 9 * (size_t)nc;
    rc = 0;
    rc |= c->d_Sd.alloc(81 * (size_t)nc);
    rc |= c->d_Cinv.alloc(81 * (size_t)nc);
    rc |= c->d_b.alloc(n9);
    rc |= c->d_x.alloc(n9);
    rc |= c->d_best.alloc(n9);
    rc |= c->d_r.alloc(n9);
    rc |= c->d_d.alloc(n9);
    rc |= c->d_Ad.alloc(n9);
    rc |= c->d_h.alloc(n9);
    rc |= c->d_h1.alloc(n9);
    rc |= c->d_s00.alloc((size_t)nc);
    rc |= c->d_scal.alloc(4);
    rc |= c->d_iscal.alloc(4);
    rc |= c->d_pcg.alloc(2);
    rc |= c->d_omega.alloc(1);
    rc |= c->d_In_bak.alloc(3 * (size_t)nc);
    rc |= c->d_ppart.alloc(3 * (size_t)nc);
    rc |= c->c_Rb.alloc(n9);
    rc |= c->c_Cb.alloc(3 * (size_t)nc);
    if (rc) return RPBA_E_NOMEM;
    if (c->fl.solver == 1) {
        const int cam0 = 0;
        rc = c->dense.plan(1, &cam0, &nc, nc);
        if (rc) return rc;
        rc = c->dense.autotune(c->st);
        if (rc) return rc;
    }
    setup_mark(c, "streams and small buffers");
    reset_state(c);
    rc = upload_tracks(c);
    if (rc) return rc;
    setup_mark(c, "points restored");
    if (c->fl.schur == 2) {
        // the product of the PCG by bytes per iteration: S formed streams
        // 648 (B + N_c) bytes (B <= min(N_c (N_c - 1) / 2, point pairs)), the
        // matrix-free product the W and Y records and the point shares,
        // (W and Y record bytes + 8) N_o + 48 N_p; the sums are the same in
        // another order
        std::vector<int> toff((size_t)c->ds.npts + 1);
        HIPCHK(hipMemcpyAsync(toff.data(), c->ds.toff, sizeof(int) * toff.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        double pairs = 0.;
        for (int i = 0; i < c->ds.npts; ++i) {
            const double t = (double)(toff[(size_t)i + 1] - toff[i]);
            pairs += t * (t - 1.) * 0.5;
        }
        const double bup = std::min(0.5 * (double)nc * (double)(nc - 1), pairs);
        const double formed = 648. * (bup + (double)nc);
        // matrix-free: both records (W and Y as their factors), the image and
        // point index of every entry, and t written and read per point
        const double esz = c->rec32 ? 4. : 8.;
        const double free_ = ((double)(2 * kWRec) * esz + 8.) * (double)c->ds.ntr + 48. * (double)c->ds.npts;
        c->implicit = c->jform || (c->fl.solver == 0 && c->fl.preconditioner == 0 && free_ < formed);
        c->ds.with_pairs = !c->implicit;
    }
    rc = upload_initial_cameras(c);
    if (rc) return rc;
    setup_mark(c, "cameras");
    // the first device-to-host copy of this size pays a one-time set-up of
    // several milliseconds: take it here rather than at the first pass boundary
    HIPCHK(hipMemcpyAsync(c->hcam, c->d_Sd.p, sizeof(double) * 15 * (size_t)nc, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(c->hb, c->d_scal.p, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return RPBA_OK;
}

extern "C" int rpba_version(void) { return RPBA_ABI_VERSION; }

// processors available to this process (omp_get_num_procs, system.h:49-76)
static int host_procs()
{
    cpu_set_t cs;
    CPU_ZERO(&cs);
    if (sched_getaffinity(0, sizeof cs, &cs) == 0) {
        const int n = CPU_COUNT(&cs);
        if (n > 0) return n;
    }
    const int n = (int)std::thread::hardware_concurrency();
    return n > 0 ? n : 1;
}

// the number of parts of CalcPLSA's partitioned branch, or 0 for the global
// solve: nrthreads = min(Threads, iround(N / 70)), Threads < 1 meaning every
// processor (rpba.cpp:85-86), partitioned when nrthreads > 1 and N > 5
// (rpbacore.cpp:79-92)
static int partition_dispatch(const rpba_flags *f, int nc)
{
    int mt = f->threads;
    if (mt < 1) mt = host_procs();
    const int minnrthreads = (int)std::lround((double)((float)nc / 70.f));  // rpbacore.cpp:79-85
    const int nrthreads = std::min(mt, minnrthreads);
    return (nrthreads > 1 && nc > 5) ? nrthreads : 0;  // rpbacore.cpp:92
}

extern "C" int rpba_ctx_create(rpba_ctx **out, rpba_block *io, const rpba_flags *flags, rpba_log_fn log, void *user)
{
    if (!out || !io || !flags || io->n_cams < 1 || io->n_pts < 0 || io->n_obs < 0) return RPBA_E_INVALID;
    if (flags->precision < 0 || flags->precision > 3 || flags->schur < 0 || flags->schur > 2) return RPBA_E_INVALID;
    // the whitened Jacobian records are read by the matrix-free product only
    if (flags->precision >= 2 && (flags->schur == 0 || flags->solver != 0 || flags->preconditioner != 0))
        return RPBA_E_INVALID;
    // the dense form factors the formed S: it needs the off-diagonal blocks
    if (flags->solver < 0 || flags->solver > 1 || (flags->solver == 1 && flags->schur == 1)) return RPBA_E_INVALID;
    // the corrected preconditioner multiplies by the formed off-diagonal blocks
    if (flags->preconditioner < 0 || flags->preconditioner > 2) return RPBA_E_INVALID;
    if (flags->preconditioner >= 1 && (flags->schur == 1 || flags->solver != 0)) return RPBA_E_INVALID;
    if (flags->preconditioner == 1 && !(flags->pc_omega >= 0.f && flags->pc_omega < 1.f)) return RPBA_E_INVALID;
    if (!io->obs_cam || !io->obs_pt || !io->obs_xy || !io->cam_P || !io->cam_f || !io->cam_d || !io->pts)
        return RPBA_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RPBA_E_HIP;
    rpba_ctx *c = new rpba_ctx();
    c->io = io;
    c->fl = *flags;
    c->log = log;
    c->user = user;
    const int rc = ctx_init(c);
    if (rc) {
        rpba_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return RPBA_OK;
}

extern "C" int rpba_ctx_run(rpba_ctx *c, int32_t max_iters, int32_t *iters_done, int32_t *finished)
{
    if (!c) return RPBA_E_INVALID;
    int done = 0;
    while (c->stage != ST_DONE && c->status == RPBA_OK && (max_iters < 0 || done < max_iters)) {
        if (c->stage == ST_SETUP) {
            const int rc = pass_setup(c);
            if (rc) {
                c->status = rc;
                break;
            }
            continue;
        }
        const int rc = lm_iteration(c);
        ++done;
        if (rc < 0) {
            c->status = rc;
            break;
        }
        if (rc == 1) {
            const int r2 = pass_end(c);
            if (r2) {
                c->status = r2;
                break;
            }
        }
    }
    if (iters_done) *iters_done = done;
    if (finished) *finished = (c->stage == ST_DONE || c->status != RPBA_OK) ? 1 : 0;
    return c->status;
}

extern "C" int rpba_ctx_reset(rpba_ctx *c)
{
    if (!c) return RPBA_E_INVALID;
    c->Pst = c->Pst_init;
    c->Ist = c->Ist_init;
    reset_state(c);
    int rc = c->ds.restore_tracks(c->st);
    if (rc) return rc;
    rc = restore_points(c);
    if (rc) return rc;
    return upload_initial_cameras(c);
}

extern "C" int rpba_ctx_download(rpba_ctx *c, rpba_stats *st)
{
    if (!c) return RPBA_E_INVALID;
    rpba_block *io = c->io;
    const int nc = c->nc;
    const int npts = c->ds.npts, ntr = c->ds.ntr;
    // kept flags in file order and the stored state's squared residuals on
    // the device; the flags, the accepted points and their original indices
    // come back (the tracks themselves stay on the device)
    unsigned char *d_ok = nullptr, *d_pk = nullptr;
    double *d_part = nullptr;
    const int g = out_groups(ntr);
    // every buffer the output kernels read must be there at its size
    if (!c->c_Pst.p || c->c_Pst.n < 12 * (size_t)nc || !c->c_Ist.p || c->c_Ist.n < 3 * (size_t)nc || !c->c_In.p ||
        c->c_In.n < 3 * (size_t)nc || (npts && (!c->d_X.p || c->d_X.n < 3 * (size_t)npts || !c->ds.nrpv)) ||
        (ntr && (!c->ds.tfo || !c->ds.tcam || !c->ds.tpt || !c->ds.ox || !c->ds.oy)))
        return RPBA_E_INVALID;
    if (hipMalloc((void **)&d_ok, std::max<size_t>((size_t)io->n_obs, 1)) != hipSuccess ||
        hipMalloc((void **)&d_pk, std::max<size_t>((size_t)io->n_pts, 1)) != hipSuccess ||
        hipMalloc((void **)&d_part, sizeof(double) * ((size_t)g + 1)) != hipSuccess) {
        if (d_ok) (void)hipFree(d_ok);
        if (d_pk) (void)hipFree(d_pk);
        if (d_part) (void)hipFree(d_part);
        return RPBA_E_NOMEM;
    }
    std::vector<double> X(3 * (size_t)npts), P(12 * (size_t)nc), In(3 * (size_t)nc);
    std::vector<int> nrpv((size_t)npts);
    double ss = 0.;
    int rc = RPBA_OK;
    if (hipMemsetAsync(d_ok, 0, (size_t)io->n_obs, c->st) != hipSuccess ||
        hipMemsetAsync(d_pk, 0, (size_t)io->n_pts, c->st) != hipSuccess)
        rc = RPBA_E_HIP;
    if (!rc) {
        launch_output(ntr, npts, c->ds.tfo, c->ds.tcam, c->ds.tpt, c->ds.nrpv, c->ds.ox, c->ds.oy, c->c_Pst.p, c->c_Ist.p,
                      c->d_X.p, d_ok, d_pk, d_part, d_part + g, c->st);
        if (hipGetLastError() != hipSuccess) rc = RPBA_E_HIP;
    }
    // the device holds the stored state from the set-up on (restore_points,
    // upload_initial_cameras); R|t from the stored state, intrinsics from the
    // last trial (OutputLSAc)
    if (!rc && ((npts && hipMemcpyAsync(X.data(), c->d_X.p, sizeof(double) * X.size(), hipMemcpyDeviceToHost, c->st) != hipSuccess) ||
                (npts && hipMemcpyAsync(nrpv.data(), c->ds.nrpv, sizeof(int) * (size_t)npts, hipMemcpyDeviceToHost, c->st) != hipSuccess) ||
                hipMemcpyAsync(P.data(), c->c_Pst.p, sizeof(double) * P.size(), hipMemcpyDeviceToHost, c->st) != hipSuccess ||
                hipMemcpyAsync(In.data(), c->c_In.p, sizeof(double) * In.size(), hipMemcpyDeviceToHost, c->st) != hipSuccess ||
                hipMemcpyAsync(&ss, d_part + g, sizeof(double), hipMemcpyDeviceToHost, c->st) != hipSuccess ||
                (io->obs_kept && io->n_obs && hipMemcpyAsync(io->obs_kept, d_ok, (size_t)io->n_obs, hipMemcpyDeviceToHost, c->st) != hipSuccess) ||
                (io->pt_kept && io->n_pts && hipMemcpyAsync(io->pt_kept, d_pk, (size_t)io->n_pts, hipMemcpyDeviceToHost, c->st) != hipSuccess)))
        rc = RPBA_E_HIP;
    if (hipStreamSynchronize(c->st) != hipSuccess && !rc) rc = RPBA_E_HIP;
    (void)hipFree(d_ok);
    (void)hipFree(d_pk);
    (void)hipFree(d_part);
    if (rc) return rc;
    for (int i = 0; i < npts; ++i) {
        const int p = nrpv[i];
        for (int j = 0; j < 3; ++j) io->pts[3 * (size_t)p + j] = (float)X[3 * (size_t)i + j];
    }
    for (int k = 0; k < nc; ++k) {
        for (int j = 0; j < 12; ++j) io->cam_P[12 * (size_t)k + j] = (float)P[12 * (size_t)k + j];
        io->cam_d[2 * (size_t)k] = (float)In[3 * (size_t)k];
        io->cam_d[2 * (size_t)k + 1] = (float)In[3 * (size_t)k + 1];
        io->cam_f[k] = (float)In[3 * (size_t)k + 2];
    }
    if (st) {
        const int64_t n = ntr;
        rpba_stats keep = *st;
        *st = c->stats;
        st->itt = c->itt;
        st->pcg_iters = c->pcg_total;
        st->sig0_final = c->sig0old;
        st->rms_px = n ? std::sqrt(ss / (double)(2 * n)) : 0.;
        st->n_obs_kept = c->ds.ntr;
        st->n_pts_kept = c->ds.npts;
        st->trace_cap = keep.trace_cap;
        st->trace_sig0 = keep.trace_sig0;
        st->trace_accept = keep.trace_accept;
        st->trace_pcg = keep.trace_pcg;
        st->trace_nrobs = keep.trace_nrobs;
        const int nt = std::min<int>((int)c->tr_sig0.size(), keep.trace_cap);
        st->trace_len = keep.trace_sig0 ? nt : 0;
        for (int j = 0; j < st->trace_len; ++j) {
            st->trace_sig0[j] = c->tr_sig0[j];
            if (st->trace_accept) st->trace_accept[j] = c->tr_acc[j];
            if (st->trace_pcg) st->trace_pcg[j] = c->tr_pcg[j];
            if (st->trace_nrobs) st->trace_nrobs[j] = c->tr_nrobs[j];
        }
    }
    return c->status;
}

extern "C" int rpba_ctx_set_timing(rpba_ctx *c, int32_t on)
{
    if (!c) return RPBA_E_INVALID;
    c->timing = on != 0;
    c->ktime.clear();
    return RPBA_OK;
}

extern "C" int rpba_ctx_tuning(rpba_ctx *c, char *buf, int32_t cap)
{
    if (!c || !buf || cap < 1) return RPBA_E_INVALID;
    const std::string t = c->fl.solver == 1 ? c->dense.tuning() : std::string("block-Jacobi PCG (no tuned kernels)");
    std::snprintf(buf, (size_t)cap, "%s", t.c_str());
    return RPBA_OK;
}

extern "C" int rpba_ctx_set_pcg_budget(rpba_ctx *c, int32_t max_iters)
{
    if (!c || max_iters < 0) return RPBA_E_INVALID;
    c->pcg_budget = max_iters;
    return RPBA_OK;
}

extern "C" int rpba_ctx_set_graph(rpba_ctx *c, int32_t on)
{
    if (!c) return RPBA_E_INVALID;
    c->pcg_graph_on = on != 0;
    return RPBA_OK;
}

extern "C" int rpba_ctx_set_lp(rpba_ctx *c, rpba_lp *model)
{
    if (!c) return RPBA_E_INVALID;
    c->lp = model;
    return RPBA_OK;
}

extern "C" int rpba_ctx_last_omega(rpba_ctx *c, float *omega)
{
    if (!c || !omega) return RPBA_E_INVALID;
    *omega = c->omega_last;
    return RPBA_OK;
}

extern "C" int rpba_ctx_lp_sample(rpba_ctx *c, int32_t n_omega, const float *omegas, int32_t *iters, float *feat,
                                  int32_t *adj_off, int32_t *adj_nbr, float *adj_w, int64_t *n_edges)
{
    if (!c || n_omega < 0 || (n_omega && (!omegas || !iters)) || !n_edges) return RPBA_E_INVALID;
    if (c->implicit || c->fl.solver != 0 || c->ds.nblk < 0 || !c->d_Su.p || c->tr_sig0.empty()) return RPBA_E_INVALID;
    LpGraphs G;
    int rc = lp_graph(c, G);
    if (rc) return rc;
    const int64_t ne = G.ne, cap = *n_edges;
    *n_edges = ne;
    HIPCHK(hipStreamSynchronize(c->st));
    if (feat) HIPCHK(hipMemcpy(feat, c->lp_feat.p, sizeof(float) * (size_t)c->nc * kLpF, hipMemcpyDeviceToHost));
    if (adj_off) HIPCHK(hipMemcpy(adj_off, c->lp_adj_off.p, sizeof(int) * ((size_t)c->nc + 1), hipMemcpyDeviceToHost));
    if (ne <= cap && ne > 0) {
        if (adj_nbr) HIPCHK(hipMemcpy(adj_nbr, c->lp_adj_nbr.p, sizeof(int) * (size_t)ne, hipMemcpyDeviceToHost));
        if (adj_w) HIPCHK(hipMemcpy(adj_w, c->lp_adj_w.p, sizeof(float) * (size_t)ne, hipMemcpyDeviceToHost));
    }
    // PCG iterations of this S and b (the last LM iteration's) per weight
    const int keep_pc = c->fl.preconditioner, keep_last = c->pcg_last;
    const float keep_w = c->omega_last;
    for (int i = 0; i < n_omega && rc == 0; ++i) {
        c->fl.preconditioner = omegas[i] > 0.f ? 1 : 0;
        PcgArgs pa = pcg_args(c);
        if (pa.omega && hipMemcpyAsync(c->d_omega.p, &omegas[i], sizeof(float), hipMemcpyHostToDevice, c->st) != hipSuccess)
            rc = RPBA_E_HIP;
        PcgScalars hs{};
        if (!rc) rc = pcg_solve(c, pa, false, hs);
        iters[i] = hs.iters;
    }
    c->fl.preconditioner = keep_pc;
    c->pcg_last = keep_last;
    c->omega_last = keep_w;
    return rc;
}

extern "C" int rpba_ctx_kernel_time(rpba_ctx *c, const char *kernel, double *total_ms, int64_t *launches)
{
    if (!c || !kernel) return RPBA_E_INVALID;
    auto it = c->ktime.find(kernel);
    if (it == c->ktime.end()) {
        if (total_ms) *total_ms = 0.;
        if (launches) *launches = 0;
        return RPBA_OK;
    }
    if (total_ms) *total_ms = it->second.first;
    if (launches) *launches = it->second.second;
    return RPBA_OK;
}

extern "C" int rpba_ctx_progress(rpba_ctx *c, int32_t *pass, int32_t *itt, double *sig0_stored)
{
    if (!c) return RPBA_E_INVALID;
    if (pass) *pass = c->it;
    if (itt) *itt = c->itt;
    if (sig0_stored) *sig0_stored = c->sig0old;
    return c->status;
}

extern "C" int rpba_ctx_sizes(rpba_ctx *c, int64_t out[6])
{
    if (!c || !out) return RPBA_E_INVALID;
    out[0] = c->nc;
    out[1] = c->ds.npts;
    out[2] = c->ds.ntr;
    out[3] = c->structure_dirty ? -1 : c->ds.npairs;
    out[4] = c->structure_dirty ? -1 : c->ds.nblk;
    out[5] = c->nrioff;
    return RPBA_OK;
}

extern "C" void rpba_ctx_destroy(rpba_ctx *c)
{
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->st2) (void)hipStreamSynchronize(c->st2);
    for (auto &e : c->pcg_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_rec) (void)hipEventDestroy(c->ev_rec);
    if (c->ev_pairs) (void)hipEventDestroy(c->ev_pairs);
    if (c->hb) (void)hipHostFree(c->hb);
    if (c->hcam) (void)hipHostFree(c->hcam);
    c->hb = nullptr;
    c->hcam = nullptr;
    for (auto &t : c->pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : c->evpool) (void)hipEventDestroy(e);
    c->ds.release();
    c->d_wx.release(); c->d_wy.release(); c->d_wxy.release(); c->d_key.release(); c->d_s00.release();
    c->d_flag.release(); c->d_flagst.release();
    c->d_X.release(); c->d_Xn.release(); c->d_X0.release(); c->d_In_bak.release(); c->d_Wrec.release(); c->d_Yrec.release(); c->d_Vinv.release(); c->d_gp.release();
    c->d_Sd.release(); c->d_Su.release(); c->d_b.release(); c->d_Cinv.release();
    c->d_x.release(); c->d_best.release(); c->d_r.release(); c->d_d.release(); c->d_Ad.release();
    c->d_h.release(); c->d_h1.release(); c->d_partial.release(); c->d_scal.release();
    c->d_ipartial.release(); c->d_iscal.release();
    c->c_Rb.release(); c->c_Cb.release(); c->c_Rn.release(); c->c_Cn.release(); c->c_In.release();
    c->c_th.release(); c->c_thn.release(); c->c_Pst.release(); c->c_Ist.release(); c->c_mask.release();
    c->d_pcg.release();
    c->d_omega.release();
    if (c->pcg_gexec) (void)hipGraphExecDestroy(c->pcg_gexec);
    if (c->pcg_graph) (void)hipGraphDestroy(c->pcg_graph);
    c->lp_nrm.release(); c->lp_eb.release(); c->lp_feat.release(); c->lp_adj_w.release();
    c->lp_adj_off.release(); c->lp_adj_nbr.release(); c->lp_goff.release();
    c->d_Dd.release();
    c->d_rowpart.release();
    c->d_colpart.release();
    c->d_dpart.release();
    c->d_t.release();
    c->d_cmpt.release(); c->d_chcam.release();
    c->d_cmox.release(); c->d_cmoy.release(); c->d_cmw.release(); c->d_cmkey.release();
    c->d_cpart.release();
    c->d_Dk.release();
    c->d_Jpm.release();
    c->d_Jcm.release();
    c->d_dmp.release();
    c->d_ptrec.release();
    c->d_ppart.release();
    if (c->st) (void)hipStreamDestroy(c->st);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    delete c;
}

// Exteval on the device (EvalLSA rpbacore.cpp:882-941): the observations are
// grouped into point-major tracks on the host (pointXXvio, file order within a
// track), uploaded, and evaluated by launch_exteval, which also forms sigma.
static int exteval(rpba_block *io, const rpba_flags *f, rpba_log_fn log, void *user, rpba_stats *stats)
{
    const int nc = io->n_cams, np = io->n_pts, no = io->n_obs;
    std::vector<int> cnt((size_t)np + 1, 0);
    for (int i = 0; i < no; ++i) {
        const int cc = io->obs_cam[i], p = io->obs_pt[i];
        if (cc < 0) continue;
        if (cc >= nc || p < 0 || p >= np) return RPBA_E_INVALID;
        cnt[(size_t)p + 1]++;
    }
    for (int i = 0; i < np; ++i) cnt[(size_t)i + 1] += cnt[i];
    const size_t nt = (size_t)cnt[np];
    std::vector<int> cam(nt);
    std::vector<float> xy(2 * nt), w(io->obs_w ? 3 * nt : 0);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < no; ++i) {
        if (io->obs_cam[i] < 0) continue;
        const size_t q = (size_t)pos[io->obs_pt[i]]++;
        cam[q] = io->obs_cam[i];
        xy[2 * q] = io->obs_xy[2 * (size_t)i];
        xy[2 * q + 1] = io->obs_xy[2 * (size_t)i + 1];
        if (io->obs_w)
            for (int j = 0; j < 3; ++j) w[3 * q + j] = io->obs_w[3 * (size_t)i + j];
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RPBA_E_HIP;
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    DBuf<int> d_off, d_cam, d_tp, d_cp, d_iw, d_ih;
    DBuf<float> d_xy, d_w, d_P, d_f, d_d, d_pts;
    DBuf<double> d_part, d_out;
    int rc = 0;
    rc |= d_off.put(cnt.data(), cnt.size(), st);
    rc |= d_cam.put(cam.data(), nt, st);
    rc |= d_xy.put(xy.data(), 2 * nt, st);
    if (io->obs_w) rc |= d_w.put(w.data(), 3 * nt, st);
    if (io->img_w) rc |= d_iw.put(io->img_w, (size_t)nc, st);
    if (io->img_h) rc |= d_ih.put(io->img_h, (size_t)nc, st);
    rc |= d_P.put(io->cam_P, 12 * (size_t)nc, st);
    rc |= d_f.put(io->cam_f, (size_t)nc, st);
    rc |= d_d.put(io->cam_d, 2 * (size_t)nc, st);
    rc |= d_pts.put(io->pts, 3 * (size_t)np, st);
    const int nb = exteval_groups(np);
    rc |= d_part.alloc((size_t)nb);
    rc |= d_tp.alloc((size_t)nb);
    rc |= d_cp.alloc((size_t)nb);
    rc |= d_out.alloc(3);
    double res[3] = {0., 0., 0.};
    if (!rc) {
        ExtArgs a;
        a.npts = np; a.ncams = nc;
        a.nraddp = (f->est_k2 != 0) + (f->est_k4 != 0) + (f->est_f != 0);
        a.toff = d_off.p; a.cam = d_cam.p; a.xy = d_xy.p; a.w = io->obs_w ? d_w.p : nullptr;
        a.P = d_P.p; a.cf = d_f.p; a.cd = d_d.p; a.pts = d_pts.p;
        a.img_w = io->img_w ? d_iw.p : nullptr; a.img_h = io->img_h ? d_ih.p : nullptr;
        a.partial = d_part.p; a.tpartial = d_tp.p; a.cpartial = d_cp.p;
        launch_exteval(a, d_out.p, st);
        if (hipGetLastError() != hipSuccess) rc = RPBA_E_HIP;
        if (hipMemcpyAsync(res, d_out.p, sizeof res, hipMemcpyDeviceToHost, st) != hipSuccess) rc = RPBA_E_HIP;
        if (hipStreamSynchronize(st) != hipSuccess) rc = RPBA_E_HIP;
    }
    d_off.release(); d_cam.release(); d_tp.release(); d_cp.release(); d_iw.release(); d_ih.release();
    d_xy.release(); d_w.release(); d_P.release(); d_f.release(); d_d.release(); d_pts.release();
    d_part.release(); d_out.release();
    (void)hipStreamDestroy(st);
    if (rc) return rc;
    const double s = res[0];
    const long long nobs = (long long)res[1], dof = (long long)res[2];
    char buf[160];
    std::snprintf(buf, sizeof buf, "ITTTT: %g %lld %lld\n", s, nobs, dof);
    if (log) log(buf, user);
    if (stats) {
        stats->exteval_sigma = s;
        stats->exteval_nobs = nobs;
        stats->exteval_dof = dof;
    }
    return RPBA_OK;
}

extern "C" int rpba_calc_plsa(rpba_block *io, const rpba_flags *flags, rpba_log_fn log, void *user, rpba_stats *stats)
{
    if (!io || !flags) return RPBA_E_INVALID;
    if (flags->exteval) return exteval(io, flags, log, user, stats);  // rpbacore.cpp:94-97
    const int nparts = partition_dispatch(flags, io->n_cams);
    if (nparts && !flags->force_global) {
        // the partitioned solve (rpbacore.cpp:100-510): plain FP64 arithmetic
        // and the block-Jacobi PCG of the reference
        if (flags->precision != 0 || flags->preconditioner != 0 || flags->schur == 1 || flags->solver != 0)
            return RPBA_E_UNSUPPORTED;
        if (!io->obs_cam || !io->obs_pt || !io->obs_xy || !io->cam_P || !io->cam_f || !io->cam_d || !io->pts ||
            io->n_cams < 1 || io->n_pts < 0 || io->n_obs < 0)
            return RPBA_E_INVALID;
        return rpba::partitioned_solve(io, flags, nparts, log, user, stats);
    }
    rpba_ctx *c = nullptr;
    int rc = rpba_ctx_create(&c, io, flags, log, user);
    if (rc) return rc;
    int done = 0, fin = 0;
    rc = rpba_ctx_run(c, -1, &done, &fin);
    const int rc2 = rpba_ctx_download(c, stats);
    rpba_ctx_destroy(c);
    return rc ? rc : rc2;
}
