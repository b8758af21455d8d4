// Kernel argument blocks and launchers of kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

namespace rpba {

// elements per entry record of Y = W V^-1 (27 used): 224 bytes in FP64, 112
// in FP32, whole 16-byte pieces; the kernels that read a wavefront's records
// read them as one contiguous span
constexpr int kRec = 28;
// elements per entry record of W = Bc^T W Bp in the matrix-free form, kept as
// its two factors: the rows of Bt = W Bc as pairs (t0_r, t1_r), r = 0..8
// (18), then the two rows of Bp (6); 192 bytes in FP64.  W[r][c] =
// t0_r Bp[0][c] + t1_r Bp[1][c], the operations that form it, so every reader
// rebuilds it bit for bit.  The matrix-free form's Y = W V^-1 is kept the
// same way as Bt^T P with P = Bp V^-1 (the Bt pairs, then P).
constexpr int kWRec = 24;

// Upper off-diagonal blocks of S (9 x 9, block list blk_k/blk_l, row-major
// order) are stored in strip slabs.  Strip s holds rows kSlab s .. kSlab s +
// kSlab - 1; its slab for column l holds the m blocks (k, l) of those rows
// that exist, k ascending, and starts at block slot base.  Inside the slab,
// element (i, j) of the slab's block number g' sits at 81 base + 9 m j +
// 9 g' + i: for a fixed column j the rows of all m blocks are contiguous, so
// lanes 9 g + i of a wavefront read them as one run of 9 m doubles.
constexpr int kSlab = 7;
// entries per point group: group c holds the points whose first entry lies in
// [kPtGroup c, kPtGroup (c + 1)).  The point kernels take a group's entries 256
// at a time; below 256 so that a group whose last point runs past the window
// still takes one pass unless that point has more than 17 entries.
constexpr int kPtGroup = 240;
// workgroups of the matrix-free product's d.S d sum (k_mf_dsum): its partials
// follow the nch chunk shares in dpart
constexpr int kMfDsumGroups = 256;
struct SuMap {
    const int *slot;       // block -> slot (position in strip, column, row order)
    const int *slot_slab;  // slot -> slab
    const int *slab_bs;    // slab -> first slot (nslab + 1)
};
__device__ __forceinline__ size_t su_at_slot(const SuMap &m, int q, int i, int j)
{
    const int s = m.slot_slab[q];
    const int base = m.slab_bs[s], n = m.slab_bs[s + 1] - base;
    return 81 * (size_t)base + (size_t)(9 * n) * j + 9 * (size_t)(q - base) + i;
}
__device__ __forceinline__ size_t su_at(const SuMap &m, int b, int i, int j) { return su_at_slot(m, m.slot[b], i, j); }

// device-resident camera state, all double, indexed by image
struct CamPtrs {
    double *Rb, *Cb;    // base rotation / centre offset of the pass (eRM, ePC)
    double *Rn, *Cn;    // trial rotation / C (eRMnew, ePCnew)
    double *In;         // trial intrinsics d0, d1, f (camerasnew)
    double *th, *thn;   // accepted / trial parameters, 9 slots (XPMatrices, XPMatricesnew)
    double *Pst;        // stored [R|t] (PMatrices)
    double *Ist;        // stored intrinsics (cameras)
    unsigned *mask;     // bit j set: slot j is estimated
};

struct PcgScalars {
    double rho, m0, mold;
    int k, done, iters, nan_flag;
};

struct LinArgs {
    int nc, npts, ntr, npgrp;
    const int *pgrp;  // point groups (first point of group c)
    double defdiff, mult1;
    CamPtrs cam;
    const int *toff, *tcam, *tpt, *camoff, *camq;
    const float *ox, *oy, *wx, *wy, *wxy;
    const double *X;
    void *Wrec, *Yrec;  // per-entry W and Y records, kRec elements, double or float (rec32)
    int rec32;
    double *Vinv, *gp, *Sd, *bvec;
    double *Dd;  // per chunk: diagonal Schur term (81) and Y g_p (9)
    double *Dk;  // per camera, matrix-free form only (else null): D_k = sum of its chunks' terms (81)
    const int *yrank;  // matrix-free form: Y records in camera-major order, entry q at yrank[q] (else null)
    int nch;     // chunks of the camera-major list (at most 64 entries of one image each)
    const int *ch_start, *ch_off;  // chunk -> first entry (nch + 1); image -> first chunk (nc + 1)
    // partition blocks (the GCP mode of LSA): per point the index of its
    // position prior or -1, the priors' centres (3 doubles each) and
    // weights (6 floats each: XX, YY, ZZ, XY, XZ, YZ); null elsewhere
    const int *gcp_of = nullptr;
    const double *gcp_c = nullptr;
    const float *gcp_w = nullptr;
    // camera-major copies of the per-entry inputs (entry j of camq: its
    // point, observation and weights), streamed by the camera kernels; null:
    // they gather through camq
    const int *cm_pt = nullptr;
    const float *cm_ox = nullptr, *cm_oy = nullptr;
    const float4 *cm_w = nullptr;  // (wx, wy, wxy, 0)
    // whitened Jacobian records (jform, precision 2): point-major and
    // camera-major copies, kJRec elements each (float when rec32), and the
    // damping of S_kk's diagonal per camera (9)
    int jform = 0;
    void *Jpm = nullptr, *Jcm = nullptr;
    double *dmp = nullptr;
    // per point, one 128-byte line for the camera side: with jform X (3),
    // the upper triangle of V^-1 (6: 00 01 02 11 12 22), g_p (3), 4 unused;
    // with ycm X (3), V^-1 (9, row-major), g_p (3), 1 unused
    double *ptrec = nullptr;
    const int4 *pg4 = nullptr;  // per point group {p0, p1, e0, e1}
    // matrix-free form with W / Y records: the point side writes W only, as
    // its factors (kWRec elements), and the point lines; k_lin_chunks forms Y
    // (as Bt^T P, kWRec elements) in camera-major order and the chunks'
    // diagonal Schur terms (needs the entry table, ptrec and ch4)
    int ycm = 0;
    const int4 *ch4 = nullptr;  // per chunk {first entry, entries, image, first chunk of the image}
};

struct SchurArgs {
    int nblk;
    const int *poff;
    const unsigned long long *pab;  // pair (a << 32 | b), cam(a) < cam(b)
    const void *Yrec, *Wrec;
    double *Su;
    SuMap sm;
    const int *sperm;  // slot -> block
    int nslab;
    const int *slab_bs;  // strip slabs (see SuMap): first slot of each slab (nslab + 1)
    long long npairs;
    int rec32;
};

struct PcgArgs {
    int nc, msize;
    const double *Sd, *Su, *Cinv, *b;
    const int *row_off, *col_off, *col_blk, *blk_k, *blk_l;
    double *x, *best, *r, *d, *h, *h1;
    double *Ad;       // S_kk d_k per camera (diagonal products of the SpMV)
    double *part;     // per camera: |alpha S d|^2, r.(h1 - h), r.h1 shares
    SuMap sm;
    // strip slabs of S (see SuMap) and the SpMV's partial products: rowpart
    // (per chunk, 7 rows x 9) and colpart (per slab, 9)
    int nslab, nchunk, nstrip, slab_chunk;
    const int *slab_bs, *slab_l, *slab_mask, *strip_s0, *chunk_off, *chunk_strip, *cs_off, *cs_list;
    double *rowpart, *colpart;
    double *dpart;    // per SpMV workgroup (or per chunk, matrix-free): its share of d.S d
    int nblk, ndpart; // upper blocks; number of dpart shares
    const double *dsum;  // what k_pcg_a adds up for d.S d: dpart, or (matrix-free) its total at dpart[nch]
    int ndsum;
    PcgScalars *S;  // two slots: launch j reads S[j & 1] and writes S[(j + 1) & 1]
    int j;          // launch index inside one PCG solve
    // matrix-free product S d = (S_kk + D_k) d_k - sum_(entries q of k) Y_q t_pt(q),
    // t_i = sum_(entries q of i) W_q^T d_cam(q), with S never formed off the diagonal
    int implicit, npts, ntr, nch, rec32, npgrp;
    const int *toff, *tcam, *tpt, *camq, *camrank, *ch_start, *ch_off;
    const int *pgrp;  // point groups (first point of group c; npgrp + 1 entries)
    const void *Wrec, *Yrec;
    const double *Dk;  // per camera: D_k = sum_q Y_q W_q^T (81)
    double *tc;        // per point: t (3)
    const int *cq_pt;  // camera-major entry -> its point
    const int *ch_cam; // chunk -> its image
    double *cpart;     // per chunk: sum_q Y_q t_pt(q) (9)
    // corrected preconditioner (S formed only): h = C r - w C N C r with the
    // weight w in device memory (null: block-Jacobi)
    float *omega = nullptr;
    // matrix-free product through the whitened Jacobian records (jform):
    // tc holds s_i = V_i^-1 t_i, Ad_k = dmp_k d_k
    int jform = 0;
    const void *Jpm = nullptr, *Jcm = nullptr;
    const double *dmp = nullptr, *Vinv = nullptr;
    const int4 *pg4 = nullptr;  // per point group {p0, p1, e0, e1}
    const int4 *ch4 = nullptr;  // per chunk {first entry, entries, image, image's first chunk}
    // block-segmented PCG: per strip, the running blocks with cameras in it;
    // the strip-slab SpMV skips the chunks of strips at 0 (null: none skipped)
    const int *strip_live = nullptr;
};

struct StepArgs {
    int nc, npts, npgrp;
    CamPtrs cam;
    const int *toff, *tcam;
    const int *pgrp;  // point groups (first point of group c)
    const void *Wrec;
    int rec32;
    const double *Vinv, *gp, *sol, *X;
    double *Xn;
    PcgScalars *S;
    // batched blocks (else null): per camera bit 0 = its block takes this
    // step (decode skipped otherwise), bit 1 = first camera of its block
    // (pose fixed, never decoded; image 0 when null)
    const unsigned char *camflag = nullptr;
    // whitened Jacobian records (point-major), or null
    int jform = 0;
    const void *Jpm = nullptr;
    const int4 *pg4 = nullptr;  // per point group {p0, p1, e0, e1}
    // 1: the W records are their factors (kWRec elements: the matrix-free
    // form with the camera side forming Y); 0: kRec-element records
    int wfac = 0;
};

struct CostArgs {
    int ntr, robust, write_key;
    double defdiff;
    CamPtrs cam;
    const int *tcam, *tpt;
    const float *ox, *oy, *wx, *wy, *wxy, *w0x, *w0y, *w0xy;
    const double *X;
    float *key;
    double *partial;
    double *terms = nullptr;  // per-entry FP64 cost terms (batched blocks), or null
    // the robust key reduced by the track-length factor of the upper median,
    // at the entry's camera-major slot (cm_key[camrank[q]]), or null
    float *cm_key = nullptr;
    const int *camrank = nullptr, *toff = nullptr;
};

struct RobustArgs {
    int nc, ntr, iflag;
    const int *camoff, *camq, *tpt, *toff, *tcam;
    const float *key, *w0x, *w0y, *w0xy;
    float *s00, *wx, *wy, *wxy;
    unsigned char *flag;
    double *partial;
    int *ipartial;
    // batched blocks (else null): iflag of the camera's block
    const int *iflag_blk = nullptr, *camblk = nullptr;
    // camera-major copies (else null): the reduced keys the median reads in
    // order, and the new weights written at the entry's camera-major slot
    const float *cm_key = nullptr;
    float4 *cm_w = nullptr;
    const int *camrank = nullptr;
};

// Exteval over tracks in point-major CSR (toff[npts + 1], entries q):
// xy[2q] = xg, w[3q] = (wx, wy, wxy) or null for unit weights.
struct ExtArgs {
    int npts, ncams, nraddp;
    const int *toff, *cam;
    const float *xy, *w, *P, *cf, *cd, *pts;
    const int *img_w, *img_h;  // null = 1 (as InputBlock sets them)
    double *partial;           // per workgroup: sum of the point terms
    int *tpartial, *cpartial;  // per workgroup: sum of t and count of points with t > 1
};

void launch_lin_points(const LinArgs &a, hipStream_t st);
void launch_schur_pairs(const SchurArgs &s, hipStream_t st);
void launch_schur_diag(const LinArgs &a, hipStream_t st);
void launch_lin_cameras(const LinArgs &a, hipStream_t st);
// whitened Jacobian records (precision 2): V, g_p and the point-major records,
// then the camera blocks S_kk, b, the damping and the camera-major records
// (no separate Schur terms)
void launch_lin_points_j(const LinArgs &a, hipStream_t st);
void launch_lin_cameras_j(const LinArgs &a, hipStream_t st);
constexpr int kJRecElems = 24;  // elements of a whitened Jacobian record
int pcg_groups(int nc);
// The camera-major entry table: for entry j of camq its point (pt), its
// observation (ox, oy) and current weights (wx, wy, wxy), copied from the
// point-major arrays (entry q to slot camrank[q]) when the structure (full)
// or the weights change, so that the camera kernels stream them; ch_cam[c]:
// the image of chunk c
struct CmTable {
    const int *camrank;
    int *pt;
    float *ox, *oy;
    float4 *w;  // (wx, wy, wxy, 0)
    int *ch_cam;
};
void launch_cm_table(const LinArgs &a, const CmTable &t, bool full, hipStream_t st);
int spmv_grid(int nblk);  // workgroups of the SpMV launch (PcgArgs::ndpart)
void launch_bj_inverse(int nc, const double *Sd, double *Cinv, hipStream_t st);
void launch_pcg_init(const PcgArgs &a, hipStream_t st);
void launch_spmv(const PcgArgs &a, hipStream_t st);

void launch_pcg_update(const PcgArgs &a, hipStream_t st);
// corrected preconditioner: after launch_pcg_init, h = C b - w C N C b, d = h,
// rho = r.h (block-Jacobi again when rho is not positive); per iteration
// launch_pcg_update_w replaces launch_pcg_update (a second product N (C r))
void launch_pcg_init_w(const PcgArgs &a, hipStream_t st);
void launch_pcg_update_w(const PcgArgs &a, hipStream_t st);
void launch_backsub(const StepArgs &a, hipStream_t st);
void launch_decode(const StepArgs &a, hipStream_t st);
int cost_blocks(int ntr);
void launch_cost(const CostArgs &a, double *out, hipStream_t st);
void launch_reweight(const RobustArgs &a, double *wsum_out, int *nflag_out, hipStream_t st);
void launch_wsum(int n, const float *wx, const float *wy, double *partial, double *out, hipStream_t st);
void launch_store_cams(const CamPtrs &cp, int nc, hipStream_t st);
// OutputLSAc on the device: obs_kept (file order) and pt_kept set to 1 for the
// kept entries and points (both cleared by the caller), *ss_out = sum of the
// squared residuals of the stored state; partial holds out_groups(ntr) doubles
int out_groups(int ntr);
void launch_output(int ntr, int npts, const int *tfo, const int *tcam, const int *tpt, const int *nrpv, const float *ox,
                   const float *oy, const double *Pst, const double *Ist, const double *X, unsigned char *obs_kept,
                   unsigned char *pt_kept, double *partial, double *ss_out, hipStream_t st);
// workgroups of the Exteval launch (partial slots needed)
int exteval_groups(int npts);
// out[0] = sigma (ITTTT value), out[1] = nobs, out[2] = dof (EvalLSA,
// rpbacore.cpp:923-938), computed on the device
void launch_exteval(const ExtArgs &a, double *out, hipStream_t st);

// batched independent blocks (one union of block-diagonal structure):
// cameras, points and track entries of block b are the contiguous ranges
// coff[b].., poff[b].., eoff[b]..
struct BatchArgs {
    int nb, nc, npts, ntr;
    const int *coff, *poff, *eoff;  // nb + 1 each
    const int *camblk, *ptblk;      // camera -> block, point -> block
    const int *tpt;                 // entry -> point
};
// per block, in entry order: out[3b] = sum terms, out[3b+1] = sum (wa + wb),
// out[3b+2] = count of flags (each input may be null); part holds
// 3 * block_slices() doubles per block
void launch_block_sums(const BatchArgs &b, const double *terms, const float *wa, const float *wb,
                       const unsigned char *flag, double *out, double *part, hipStream_t st);
// per block: 1 if a camera step (as float) or a trial point is not finite;
// part holds block_slices() ints per block
void launch_block_nan(const BatchArgs &b, const double *sol, const double *Xn, int *nanb, int *part, hipStream_t st);
int block_slices();
// StoreResultsIteration for the blocks with accept[b] != 0: cameras, points, flags
void launch_batch_store(const BatchArgs &b, const unsigned char *accept, const CamPtrs &cp, const double *Xn,
                        double *X, const unsigned char *flag, unsigned char *flagst, hipStream_t st);
// trial intrinsics back from the copy taken before the step, for blocks with nanb[b]
void launch_batch_restore_in(const BatchArgs &b, const int *nanb, const double *In_bak, double *In, hipStream_t st);
void launch_copy(double *dst, const double *src, size_t n, hipStream_t st);
// the blocks' kept-point and entry ranges after a structure build: ptblk per
// kept point, poff / eoff (nb + 1 each) from the kept points' union indices
// nrpv (ascending) and the blocks' union point offsets uptoff
void launch_batch_ranges(int npts, int nb, const int *nrpv, const int *uptoff, const int *toff, int *ptblk, int *poff,
                         int *eoff, hipStream_t st);

}  // namespace rpba
