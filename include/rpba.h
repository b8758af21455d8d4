/*
 * rpba.h -- C ABI of the MI355X robust Levenberg-Marquardt Schur-complement
 * bundle adjustment (global "rbacore" solve of RPBA on one GPU).
 *
 * Every entry point takes plain pointers and sizes; the caller owns all host
 * buffers, the library owns device memory for the duration of a call (or of a
 * context).  Outputs are written in place, as the reference overwrites its
 * arguments.
 *
 * Reference interface replaced (helmayer/RPBA):
 *   rpba_calc_plsa   <- void CalcPLSA(MatrixXf& XX, vector<MatrixXf>& xg,
 *                        wx, wy, wxy, vector<vector<IMAPNR>>& pointXXvio,
 *                        vector<PMat>& PMatrices, vector<Camera>& cameras,
 *                        iwidth, iheight, bool robustflag,
 *                        vector<bool>& addflagsin, ostream& os, int mt,
 *                        bool extevalflag)
 *                       (rpba.h:693-703, rpbacore.cpp:28-531)
 *   rpba_ctx_*       <- the same solve split into create / run / download so a
 *                       caller can time a fixed number of LM iterations
 *                       (LSA::LSA rbacore.cpp:37-123, LSA::Adjust 333-616,
 *                       LSA::OutputLSAc 644-666)
 *   rpba_read_bal / rpba_write_bal / rpba_read_parameters
 *                    <- InputBlock / OutputBlock (rpba.cpp:180-338) and the
 *                       `parameters` INI reader of main (rpba.cpp:53-80)
 */
#ifndef RPBA_MI355X_H
#define RPBA_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPBA_ABI_VERSION 1

/* Return codes.  The reference returns void and prints these conditions; the
 * same text lines go to the log callback and the last stored state is still
 * written back, as rpbacore.cpp:523 does after an early return from Adjust. */
enum rpba_status {
    RPBA_OK = 0,
    RPBA_E_NO_OBS_IN_IMAGE = -1, /* "BA-Error: No observations in image" rbacore.cpp:414-425 */
    RPBA_E_LOW_REDUNDANCY = -2,  /* "BA-Error: Low redundancy"           rbacore.cpp:468-476 */
    RPBA_E_NAN_RESULT = -3,      /* "BA-Error: NANs in result vector"    rbacore.cpp:541-546 */
    RPBA_E_INVALID = -10,        /* malformed arguments (bad index, duplicate camera in a track) */
    RPBA_E_UNSUPPORTED = -11,    /* a form selector the partitioned solve does not take (it runs plain FP64, block-Jacobi PCG) */
    RPBA_E_SCRATCH = -12,        /* input larger than a no-allocation entry point's fixed scratch */
    RPBA_E_BUDGET = -13,         /* a PCG solve reached the bound of rpba_ctx_set_pcg_budget */
    RPBA_E_HIP = -20,            /* HIP runtime error */
    RPBA_E_NOMEM = -21           /* host or device allocation failed */
};

/* One bundle block in flat structure-of-arrays form.
 *
 * Observations are listed in file order.  The reference's per-image lists
 * (xg[image], wx[image] ...) and tracks pointXXvio[point] = {(image, pointnr)}
 * are exactly what InputBlock builds from such a list (rpba.cpp:219-238):
 * pointnr is the rank of the observation among the observations of its image,
 * and a track lists its observations in file order. */
typedef struct rpba_block {
    int32_t n_cams;          /* images (= cameras; one camera per image)       */
    int32_t n_pts;           /* points                                         */
    int32_t n_obs;           /* observations                                   */
    const int32_t *obs_cam;  /* [n_obs] image index; < 0 = absent (IMAPNR.image = -1) */
    const int32_t *obs_pt;   /* [n_obs] point index                            */
    const float *obs_xy;     /* [2*n_obs] xg: (-x_file, -y_file) per observation */
    const float *obs_w;      /* [3*n_obs] (wx, wy, wxy) or NULL = unit weights */
    float *cam_P;            /* [12*n_cams] row-major 3x4 [R|t] (PMat), in/out */
    float *cam_f;            /* [n_cams] K(0,0) = K(1,1), in/out               */
    float *cam_d;            /* [2*n_cams] distpara[0..1] (k24out-scaled), in/out */
    float *pts;              /* [3*n_pts] point XYZ (XX rows 0..2), in/out     */
    const int32_t *img_w;    /* [n_cams] image widths or NULL (= 1, as InputBlock sets) */
    const int32_t *img_h;    /* [n_cams] image heights or NULL (= 1)           */
    uint8_t *obs_kept;       /* [n_obs] out (may be NULL): 1 = still in its track */
    uint8_t *pt_kept;        /* [n_pts] out (may be NULL): 1 = point kept      */
} rpba_block;

/* The `parameters` controls (rpba.cpp:56-64) plus the form selectors. */
typedef struct rpba_flags {
    int32_t robust;          /* Robust  */
    int32_t exteval;         /* Exteval */
    int32_t threads;         /* Threads (< 1: all host cores, rpba.cpp:88-89)   */
    int32_t est_k2;          /* k2 -> addflags[0] */
    int32_t est_k4;          /* k4 -> addflags[1] */
    int32_t est_f;           /* f  -> addflags[2] */
    int32_t k24out;          /* k24out (I/O scaling only, rpba.cpp:255-262)     */
    int32_t precision;       /* 0 = plain FP64; 1 = FP32 Jacobians and W V^-1 W^T
                                products with FP64 sums, S, PCG, residuals and cost;
                                2 = whitened Jacobians L^T Bc, L^T Bp (observation
                                weight W = L L^T) rounded to FP32 once, and U, V,
                                W, g, the Schur terms and the PCG product formed
                                from them in FP64 (S stays positive definite at
                                every rounding); 3 = the same records in FP64.
                                2 and 3 take the matrix-free product (schur 1 or
                                2), block-Jacobi and solver 0 only            */
    int32_t preconditioner;  /* 0 = block-Jacobi C = S_kk^-1 (rbacore.cpp:1838-1846);
                                1 = block-Jacobi with the first-order Neumann
                                correction C - w C N C (N: off-diagonal part of S),
                                w = pc_omega;
                                2 = the same with w predicted per LM iteration by
                                the learned graph network (rpba_lp_*), block-Jacobi
                                when it declines or the PCG breaks down      */
    int32_t force_global;    /* 1 = run the global LSA even where the reference
                                would dispatch to the partitioned path          */
    int32_t schur;           /* 0 = S formed block by block (the reference's
                                profile store); 1 = matrix-free: only the diagonal
                                blocks are formed and the PCG multiplies through
                                W and Y = W V^-1 (no covisibility pairs built);
                                2 = whichever streams fewer bytes per PCG
                                iteration: 648 (B + N_c) with B bounded by
                                min(N_c (N_c - 1) / 2, point pairs), against
                                (W and Y record bytes + 8) N_o + 48 N_p (formed with
                                solver 1 or a corrected preconditioner)      */
    int32_t solver;          /* 0 = block-Jacobi PCG (SparsePCG, rbacore.cpp:1807-1914);
                                1 = dense Cholesky of the reduced camera system on
                                the FP64 matrix cores (exact solve; a form of its
                                own, for small or dense S)                       */
    float pc_omega;          /* preconditioner 1: the Neumann weight w in [0, 1) */
} rpba_flags;

/* Optional per-solve report.  sigma values are sqrt(sig0 * ihw2) exactly as the
 * "Sigma0:" log line prints them (rbacore.cpp:583-597). */
typedef struct rpba_stats {
    double sigma_in[2];      /* per pass: first value of the "Sigma0:" line    */
    double sigma_out[2];     /* per pass: second value (best stored)           */
    int32_t redundancy[2];
    int32_t nobs[2];
    int32_t nunknowns[2];
    int32_t passes_done;     /* 0, 1 or 2 */
    int32_t itt;             /* LM iterations (the reference's itt counter)    */
    int64_t pcg_iters;       /* total PCG iterations                           */
    double sig0_final;       /* sig0 of the stored state (sigma0^2 / ihw2)     */
    double rms_px;           /* sqrt(sum |r|^2 / 2n) over kept observations, stored state */
    int32_t n_obs_kept;
    int32_t n_pts_kept;
    double exteval_sigma;    /* ITTTT value (rpbacore.cpp:938)                 */
    int64_t exteval_nobs;
    int64_t exteval_dof;
    /* per-iteration trace (caller-owned arrays of trace_cap entries, or NULL) */
    int32_t trace_cap;
    int32_t trace_len;
    double *trace_sig0;      /* sig0 of each trial (after the wsumg/redundancy scaling) */
    int32_t *trace_accept;   /* 1 = StoreResultsIteration ran                  */
    int32_t *trace_pcg;      /* PCG iterations of that LM iteration            */
    int32_t *trace_nrobs;    /* nrobs used for that iteration's redundancy     */
} rpba_stats;

typedef void (*rpba_log_fn)(const char *line, void *user);

/* Whole solve, device resident.  Mirrors CalcPLSA: rpbacore.cpp:28-99
 * (setup, dispatch, Exteval), 511-527 (single LSA: ctor, Adjust(robust, os,
 * false, -1), OutputLSAc) and, where the reference dispatches to it
 * (nrthreads = min(Threads, iround(N / 70)) > 1 and N > 5, Threads < 1 = every
 * processor of the host, rpbacore.cpp:79-92), the partitioned solve
 * (rpbacore.cpp:100-510): per outer iteration one GCP-mode LSA step of every
 * image part and the per-point re-triangulation CalcLSAXMulti, with the image
 * partition from rpba_partition_images.  Its stats: itt = outer iterations,
 * trace_* per outer iteration (sig0g, 1 when saved as the best, the PCG
 * iterations of its parts, nrobs), sig0_final = the best sig0g, rms_px of the
 * output state; it takes plain FP64 forms only (RPBA_E_UNSUPPORTED
 * otherwise).  force_global runs the global LSA instead. */
int rpba_calc_plsa(rpba_block *io, const rpba_flags *flags,
                   rpba_log_fn log, void *user, rpba_stats *stats);

/* The image partition of the partitioned solve; replaces PartitionpointXXv
 * (rpbacore.cpp:945-1037), which calls METIS_PartGraphRecursive on the
 * covisibility graph (images joined when they share a point of two or more
 * observations; vertex weight (int) pow(sum of the track lengths of the
 * image's points, 0.3333)).  Here: deterministic recursive bisection of the
 * same graph with the same weights (graph growing from a pseudo-peripheral
 * image, boundary moves that cut fewer edges within a 3% balance), not
 * METIS's.  part[n_cams] receives part ids 0 .. n_parts - 1, every part
 * non-empty; 1 <= n_parts <= n_cams. */
int rpba_partition_images(int32_t n_cams, int32_t n_obs, const int32_t *obs_cam, const int32_t *obs_pt,
                          int32_t n_pts, int32_t n_parts, int32_t *part);

/* Resumable form of the same global solve. */
typedef struct rpba_ctx rpba_ctx;
/* LSA::LSA + uploads (rbacore.cpp:37-123). The block must stay valid until
 * rpba_ctx_destroy (download writes back into it). */
int rpba_ctx_create(rpba_ctx **ctx, rpba_block *io, const rpba_flags *flags,
                    rpba_log_fn log, void *user);
/* Continue LSA::Adjust for at most max_iters LM iterations (max_iters < 0: to
 * the end).  *iters_done receives the iterations run, *finished 1 when both
 * passes are complete (or the solve stopped on an error, returned as status). */
int rpba_ctx_run(rpba_ctx *ctx, int32_t max_iters, int32_t *iters_done, int32_t *finished);
/* Restore the state right after rpba_ctx_create (re-uploads the initial
 * tracks and cameras from the context's host copies). */
int rpba_ctx_reset(rpba_ctx *ctx);
/* OutputLSAc (rbacore.cpp:644-666): write the stored state back to the block. */
int rpba_ctx_download(rpba_ctx *ctx, rpba_stats *stats);
/* Bound every PCG solve of the context to max_iters iterations (0, the
 * default: the reference's cap max(1000, n), rbacore.cpp:1848-1849); a PCG
 * that reaches the bound stops the run with RPBA_E_BUDGET, the stored state
 * kept.  No counterpart in the reference: for callers that evaluate forms
 * under a time budget. */
int rpba_ctx_set_pcg_budget(rpba_ctx *ctx, int32_t max_iters);
/* 1: PCG iterations are replayed two at a time from a hipGraph; 0 (default):
 * launched one by one, which measured as fast.  Same results either way. */
int rpba_ctx_set_graph(rpba_ctx *ctx, int32_t on);
/* Kernel timing of the last rpba_ctx_run with timing on: name -> total ms and
 * launches, measured with HIP events on the context's stream. */
int rpba_ctx_set_timing(rpba_ctx *ctx, int32_t on);
int rpba_ctx_kernel_time(rpba_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches);
/* Dense form (solver = 1): the trailing-update variant the autotuner chose at
 * rpba_ctx_create and the timings it compared, one line.  No counterpart in
 * the reference (its solve has no tuned kernels). */
int rpba_ctx_tuning(rpba_ctx *ctx, char *buf, int32_t cap);
/* Progress of the running solve: pass (0, 1; 2 = done), LM iterations (itt)
 * and sig0 of the stored state (sig0old, rbacore.cpp:568) in the current pass. */
int rpba_ctx_progress(rpba_ctx *ctx, int32_t *pass, int32_t *itt, double *sig0_stored);
/* Problem sizes after the point filter of PointXXvIn (rbacore.cpp:1954-1955)
 * and the current covisibility structure: out[0..5] = cameras, points,
 * observations, point pairs, off-diagonal S blocks, S dimension (nrioff). */
int rpba_ctx_sizes(rpba_ctx *ctx, int64_t out[6]);
void rpba_ctx_destroy(rpba_ctx *ctx);

/* Batched dense solve on the FP64 matrix cores: nsys independent symmetric
 * positive definite systems S_m x_m = b_m of dimension 9 * ncam[m], each S_m
 * full and row-major, concatenated in S; b and x concatenated likewise.  The
 * kernels are those of the dense form (solver = 1).  variant < 0: the
 * autotuner picks the trailing-update variant; else variant 0 (LDS-staged) or
 * 1 (direct) with a persistent grid of `cap` workgroups.  Every variant gives
 * the same bits.  No counterpart in the reference (its systems are solved by
 * SparsePCG, rbacore.cpp:1807-1914); exposed for callers with their own
 * batches of camera systems. */
int rpba_dense_solve_batch(int32_t nsys, const int32_t *ncam, const double *S, const double *b, double *x,
                           int32_t variant, int32_t cap);

/* Batched independent blocks: nblocks bundle blocks solved together on one
 * GPU, each by the global robust LM of LSA::Adjust (rbacore.cpp:333-616) with
 * its own LM state, as the partitioned mode solves its METIS blocks one LSA
 * per OpenMP thread (rpbacore.cpp:219-235).  Every LM iteration of all blocks
 * runs as one launch sequence over the union of the blocks (replayed from
 * hipGraphs).  The blocks' reduced camera systems are solved by the
 * reference's block-Jacobi SparsePCG segmented by block (solver = 0: one SpMV
 * over the union per PCG iteration, each block with its own scalars, best
 * iterate, stopping rule and cap max(1000, nrioff)) or factored together by
 * the batched dense Cholesky (solver = 1).  schur must be 0, precision 0 or 1,
 * preconditioner 0.  Passes run in lockstep: a block that ends a pass early
 * waits, unchanged.  Each block needs at least 3 images, and all blocks the
 * same image sizes.
 * Block results are written back into the blocks by rpba_batch_download;
 * stats and status are arrays of nblocks entries (either may be NULL). */
typedef struct rpba_batch rpba_batch;
typedef void (*rpba_batch_log_fn)(int32_t block, const char *line, void *user);
int rpba_batch_create(rpba_batch **batch, int32_t nblocks, rpba_block *blocks, const rpba_flags *flags,
                      rpba_batch_log_fn log, void *user);
/* at most max_iters lockstep LM iterations (< 0: to the end) */
int rpba_batch_run(rpba_batch *batch, int32_t max_iters, int32_t *iters_done, int32_t *finished);
int rpba_batch_reset(rpba_batch *batch);
int rpba_batch_download(rpba_batch *batch, rpba_stats *stats, int32_t *status);
/* 1 (default): replay the LM iteration from a hipGraph; 0: eager launches */
int rpba_batch_set_graph(rpba_batch *batch, int32_t on);
int rpba_batch_set_timing(rpba_batch *batch, int32_t on);
int rpba_batch_kernel_time(rpba_batch *batch, const char *kernel, double *total_ms, int64_t *launches);
/* solver 0: the algorithmic HBM bytes of the SpMV launches timed since
 * rpba_batch_set_timing(1): per launch, 648 bytes for each upper block of S of
 * the blocks whose PCG still runs and for every diagonal block, 72 + 72 + 8
 * bytes per camera (d in, the diagonal product out, the row partials) */
int rpba_batch_spmv_bytes(rpba_batch *batch, double *bytes);
int rpba_batch_tuning(rpba_batch *batch, char *buf, int32_t cap);
/* out[0..5] = cameras, points, observations, pairs, off-diagonal S blocks,
 * doubles of dense storage, over the union */
int rpba_batch_sizes(rpba_batch *batch, int64_t out[6]);
void rpba_batch_destroy(rpba_batch *batch);
/* create + run + download + destroy */
int rpba_calc_plsa_batch(int32_t nblocks, rpba_block *blocks, const rpba_flags *flags, rpba_batch_log_fn log,
                         void *user, rpba_stats *stats, int32_t *status);

/* Learned correction of the block-Jacobi preconditioner (preconditioner = 2).
 * A graph network over the camera covisibility graph of S (node per camera,
 * edge per non-zero block S_kl with weight |S_kl|_F / sqrt(|S_kk|_F |S_ll|_F),
 * RPBA_LP_FEATURES node features from the 9x9 block S_kk and the node's
 * edges; 3 message-passing layers of width 64, BF16 weights and activations,
 * FP32 arithmetic) predicts the weight w of M^-1 = C - w C N C.  Trained by
 * the library's own kernels (Adam on (w - w*)^2, w* the weight with the
 * fewest PCG iterations on a logged S).  No counterpart in the reference (its
 * SparsePCG uses block-Jacobi, rbacore.cpp:1838-1846); the PCG stopping rule
 * is unchanged. */
#define RPBA_LP_FEATURES 48
typedef struct rpba_lp rpba_lp;
/* graphs in CSR over global node indices; the adjacency of every graph must
 * be symmetric with equal weights both ways (as the camera graph of S is) */
typedef struct rpba_lp_graphs {
    int32_t n_graphs, n_nodes;
    int64_t n_edges;
    const int32_t *goff;     /* [n_graphs + 1] first node of each graph */
    const float *feat;       /* [n_nodes * RPBA_LP_FEATURES] */
    const int32_t *adj_off;  /* [n_nodes + 1] */
    const int32_t *adj_nbr;  /* [n_edges] */
    const float *adj_w;      /* [n_edges] */
    const float *label;      /* [n_graphs] target weights, or NULL */
} rpba_lp_graphs;
/* params NULL: random initialisation from seed */
int rpba_lp_create(rpba_lp **model, const float *params, uint64_t seed);
void rpba_lp_destroy(rpba_lp *model);
int rpba_lp_n_params(void);
int rpba_lp_get_params(rpba_lp *model, float *params);
int rpba_lp_predict(rpba_lp *model, const rpba_lp_graphs *graphs, float *omega);
/* loss = mean over graphs of (w - label)^2 and its gradient (rpba_lp_n_params floats) */
int rpba_lp_gradient(rpba_lp *model, const rpba_lp_graphs *graphs, float *omega, float *loss, float *grad);
/* full-batch Adam steps; loss_trace[steps] (may be NULL) */
int rpba_lp_train(rpba_lp *model, const rpba_lp_graphs *graphs, int32_t steps, float lr, float *loss_trace);
/* attach a model to a context (preconditioner = 2; NULL: block-Jacobi) */
int rpba_ctx_set_lp(rpba_ctx *ctx, rpba_lp *model);
/* the weight w the last PCG of the context ran with */
int rpba_ctx_last_omega(rpba_ctx *ctx, float *omega);
/* training sample from the context's last LM iteration (S formed, PCG
 * solver): the camera graph of its S (feat [n_cams * RPBA_LP_FEATURES],
 * adj_off [n_cams + 1], adj_nbr / adj_w [*n_edges capacity in, count out;
 * written when the capacity suffices]) and the PCG iterations of that S and b
 * for each weight in omegas (0 = block-Jacobi) */
int rpba_ctx_lp_sample(rpba_ctx *ctx, int32_t n_omega, const float *omegas, int32_t *iters, float *feat,
                       int32_t *adj_off, int32_t *adj_nbr, float *adj_w, int64_t *n_edges);

/* Exteval (EvalLSA, rpbacore.cpp:882-941) from device-resident tracks, with
 * no device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same kernels as
 * rpba_calc_plsa with exteval = 1 run on `stream` (a hipStream_t; NULL = the
 * null stream) and the call returns without synchronising.
 *   toff [n_pts + 1]: the tracks (pointXXvio) as CSR offsets over entries q;
 *   tcam [nt], txy [2 nt] (xg), tw [3 nt] (wx, wy, wxy) or NULL = unit weights;
 *   cam_P [12 n_cams], cam_f [n_cams], cam_d [2 n_cams], pts [3 n_pts];
 *   img_w / img_h [n_cams] or NULL (= 1);
 *   out [3] (device): sigma (the ITTTT value), nobs, dof.
 * All pointers are device pointers.  Partial sums go to a fixed device array
 * of the library (one per device, held until the process ends), which limits
 * n_pts to rpba_exteval_device_max_points() (8,388,608); larger inputs return
 * RPBA_E_SCRATCH.  Calls that overlap on one device race on that array. */
int rpba_exteval_device(int32_t n_cams, int32_t n_pts, const int32_t *toff, const int32_t *tcam,
                        const float *txy, const float *tw, const float *cam_P, const float *cam_f,
                        const float *cam_d, const float *pts, const int32_t *img_w, const int32_t *img_h,
                        int32_t est_k2, int32_t est_k4, int32_t est_f, double *out, void *stream);
int rpba_exteval_device_max_points(void);

/* BAL text I/O (rpba.cpp:180-338), host only.
 * rpba_read_bal allocates the arrays of *io with malloc; free with
 * rpba_free_block.  iwidth = iheight = 1 as InputBlock sets them. */
int rpba_read_bal(const char *path, int32_t k24out, rpba_block *io);
int rpba_write_bal(const char *path, const rpba_block *io, int32_t k24out);
void rpba_free_block(rpba_block *io);
/* `parameters` INI (rpba.cpp:53-80): fills flags, input dir and file name. */
int rpba_read_parameters(const char *path, rpba_flags *flags,
                         char *input_dir, int32_t dir_cap,
                         char *file, int32_t file_cap);

/* Library version and device info for logs: returns RPBA_ABI_VERSION. */
int rpba_version(void);

#ifdef __cplusplus
}
#endif
#endif /* RPBA_MI355X_H */
