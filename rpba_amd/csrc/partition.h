// The partitioned RPBA solve (CalcPLSA's parallel branch, rpbacore.cpp:100-510)
// on one GPU: the image partition, the per-point re-triangulation kernel
// (CalcLSAXMulti, rpbacore.cpp:536-773) and the block-segmented PCG that
// solves every part's reduced camera system in one launch sequence.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/rpba.h"
#include "kernels.h"

namespace rpba {

// CalcLSAXMulti for every point with a track, one thread per point.  Tracks
// are per-point segments trk[base[i] .. base[i] + len[i]) of file-order
// observation indices (compacted in place by robust deletion); the parts a
// point's initial track touches are its slots soff[i] .. soff[i+1] (part id
// spart, pitflag pit, stored information w3: 6 floats, and a 9-double
// workspace nv).  Per point out: nout (track length left, 0 = cleared), sg
// (its share of sig0g, the float sig0opt / wsum widened), rd (its share of the
// redundancy).  keys (robustmode 1): the robust key of each entry, by
// position.  wk, dl: per-entry workspace (current weights, deletion flags).
struct LsaxArgs {
    int npts, P, robustmode, weightonly, wxyflag;
    const long long *base;
    const int *len;
    int *trk;
    double *X;
    const double *PMd;
    const float *cf, *cd;
    const float *oxy, *ow;  // 2 per observation; 3 per observation or null (unit)
    const int *obs_cam, *part;
    const float *sig0wv;
    const long long *soff;
    const int *spart;
    const unsigned char *pit;
    float *w3;
    double *nv;
    float *wk;
    unsigned char *dl;
    float *keys;
    int *nout;
    double *sg;
    int *rd;
};
void launch_lsax(const LsaxArgs &a, hipStream_t st);

// per-block PCG scalars of the segmented solve
struct SegScal {
    double rho, m0, mold;
    int k, done, iters, msize;
};

// SparsePCG (rbacore.cpp:1807-1914) for many independent blocks of the union
// S at once: the union S d by the strip-slab SpMV (launch_spmv), then per
// camera the gather of (S d)_k (k_seg_gather) and per block the update with
// its own alpha, beta, best iterate and stopping rule (k_seg_step).  Cameras
// of block b are boff[b] .. boff[b+1]; cblk maps a camera to its block.
struct SegArgs {
    int nc, nb;
    const int *boff, *cblk;
    SegScal *S;
    int *ndone;     // blocks that stopped (integer counter)
    double *Sdf;    // (S d)_k gathered per camera (9 each)
    double *e;      // d_k . (S d)_k per camera
    // per strip of kSlab cameras, the running blocks with cameras in it (the
    // SpMV skips strips at 0; PcgArgs::strip_live points here), or null
    int *strip_live = nullptr;
    int nstrip = 0;
};
void launch_seg_init(const PcgArgs &a, const SegArgs &s, hipStream_t st);
void launch_seg_iter(const PcgArgs &a, const SegArgs &s, hipStream_t st);
// the iteration without its SpMV: the gather of (S d)_k and the block updates
void launch_seg_update(const PcgArgs &a, const SegArgs &s, hipStream_t st);
// per block: 1 when the float camera step or a point step of the block is
// NaN or infinite (the reference's check, rbacore.cpp:537-546); flag[nb]
// cleared by the launcher
void launch_seg_nan(int nc, int npts, const int *cblk, const int *ptblk, const double *sol, const double *Xn,
                    int nb, int *flag, hipStream_t st);

// the image partition that replaces METIS_PartGraphRecursive
// (rpbacore.cpp:945-1037): recursive bisection of the covisibility graph
// (images joined when they share a point), vertex weight
// (int) pow(sum of the track lengths of the image's points, 0.3333)
int partition_images(int nc, int nobs, const int *obs_cam, const int *obs_pt, int npts, int nparts, int *part);

// CalcPLSA's partitioned solve on the flat block (P parts)
int partitioned_solve(rpba_block *io, const rpba_flags *f, int P, rpba_log_fn log, void *user, rpba_stats *st);

}  // namespace rpba
