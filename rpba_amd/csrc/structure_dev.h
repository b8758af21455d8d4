// Device-resident index structures of one bundle block.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/rpba.h"
#include "kernels.h"
#include "structure.h"

namespace rpba {

// point-major track arrays (one set; RobustDelete compacts into a second set
// and swaps)
struct Tracks {
    int *toff = nullptr, *nrpv = nullptr, *tcam = nullptr, *tpt = nullptr, *tfo = nullptr;
    float *ox = nullptr, *oy = nullptr, *w0x = nullptr, *w0y = nullptr, *w0xy = nullptr;
    double *X = nullptr;
    int alloc(size_t np, size_t nt);
    void release();
};

struct DevStruct {
    int nc = 0, npts = 0, ntr = 0, nblk = 0;
    long long npairs = 0;
    size_t sizes_cap_pts = 0, sizes_cap_tr = 0;
    // tracks: entries of point i are [toff[i], toff[i+1]); nrpv: original point
    // index; tfo: file-order observation index; w0*: the original weights
    int *toff = nullptr, *nrpv = nullptr, *tcam = nullptr, *tpt = nullptr, *tfo = nullptr;
    float *ox = nullptr, *oy = nullptr, *w0x = nullptr, *w0y = nullptr, *w0xy = nullptr;
    // camera-major order: entries of image k are camq[camoff[k] .. camoff[k+1])
    int *camoff = nullptr, *camq = nullptr;
    size_t camoff_cap = 0, camq_cap = 0;
    int *camrank = nullptr;  // entry -> its position in camera-major order (camrank[camq[j]] = j)
    size_t camrank_cap = 0;
    // point groups of the point kernels: group c holds the points whose first
    // entry lies in [kPtGroup c, kPtGroup (c + 1)): points pgrp[c] .. pgrp[c+1]-1
    int *pgrp = nullptr;
    size_t pgrp_cap = 0;
    int npgrp = 0;
    // per point group {first point, end point, first entry, end entry}: one
    // load gives a point kernel its group
    int4 *pg4 = nullptr;
    size_t pg4_cap = 0;
    // chunks of at most kChunk entries of one image over camq: ch_start (nch + 1
    // entry offsets), ch_off (nc + 1 first chunk of each image)
    static constexpr int kChunk = 64;
    int nch = 0;
    int *ch_start = nullptr, *ch_off = nullptr;
    size_t ch_start_cap = 0, ch_off_cap = 0;
    // per chunk {first entry, entries, image, 1 if the image's first chunk}
    int4 *ch4 = nullptr;
    size_t ch4_cap = 0;
    // pairs (a << 32 | b), cam(a) < cam(b), grouped by block in row-major
    // block order; poff: pairs of block b; bk/bl: block (k, l); roff: blocks of
    // row k; coff/cblk: blocks of column l (ids, k ascending)
    unsigned long long *pab = nullptr;
    int *poff = nullptr, *bk = nullptr, *bl = nullptr, *roff = nullptr, *coff = nullptr, *cblk = nullptr;
    size_t pab_cap = 0, poff_cap = 0, bk_cap = 0, bl_cap = 0, roff_cap = 0, coff_cap = 0, cblk_cap = 0;
    // strip slabs (the storage order of the upper blocks of S): strip s holds
    // rows kSlab s .. kSlab s + kSlab - 1; its slab for column l holds the
    // blocks (k, l) of those rows, k ascending.  sslot: block -> slot (its
    // position in strip, column, row order), sperm the inverse; slot_slab:
    // slot -> slab; slab_bs (nslab + 1): first slot of each slab; slab_l,
    // slab_mask: the slab's column and the rows k - kSlab s present (bits);
    // strip_s0 (nstrip + 1): first slab of each strip.  Chunks of at most
    // slab_chunk consecutive slabs of one strip: chunk_off (nstrip + 1): first
    // chunk of each strip; chunk_strip: the strip of each chunk.  cs_off
    // (nc + 1) / cs_list: the slabs of column l in strip order.
    int nslab = 0, nstrip = 0, nchunk = 0, slab_chunk = 8;
    int spmv_waves = 0;  // wavefronts of the SpMV launch (sets the chunk length)
    int *sslot = nullptr, *sperm = nullptr, *slot_slab = nullptr, *slab_bs = nullptr, *slab_l = nullptr;
    int *slab_mask = nullptr, *strip_s0 = nullptr, *chunk_off = nullptr, *chunk_strip = nullptr;
    int *cs_off = nullptr, *cs_list = nullptr;
    unsigned *slab_key = nullptr, *slab_cnt = nullptr;
    size_t sslot_cap = 0, sperm_cap = 0, slot_slab_cap = 0, slab_bs_cap = 0, slab_l_cap = 0, slab_mask_cap = 0;
    size_t strip_s0_cap = 0, chunk_off_cap = 0, chunk_strip_cap = 0, cs_off_cap = 0, cs_list_cap = 0;
    size_t slab_key_cap = 0, slab_cnt_cap = 0;
    int strip_slabs(hipStream_t st);
    SuMap sumap() const { return SuMap{sslot, slot_slab, slab_bs}; }
    // scratch
    unsigned *u32a = nullptr, *key_in = nullptr, *key_out = nullptr, *ukey = nullptr, *ucnt = nullptr, *nruns = nullptr;
    unsigned *key2 = nullptr, *key2s = nullptr;
    unsigned long long *val_in = nullptr;
    int *i32a = nullptr, *i32b = nullptr, *ppoff = nullptr, *bid = nullptr, *newi = nullptr, *newoff = nullptr;
    size_t u32a_cap = 0, key_in_cap = 0, key_out_cap = 0, ukey_cap = 0, ucnt_cap = 0, nruns_cap = 0;
    size_t key2_cap = 0, key2s_cap = 0, val_in_cap = 0, i32a_cap = 0, i32b_cap = 0, ppoff_cap = 0, bid_cap = 0;
    size_t newi_cap = 0, newoff_cap = 0;
    unsigned char *tmp = nullptr;
    size_t tmp_cap = 0;

    Tracks init;           // snapshot of the initial tracks (restore_tracks)
    Tracks spare;          // RobustDelete's output buffers, swapped with the live tracks
    // RobustDelete updates the pairs and the camera-major order in place of a
    // rebuild: emap (old entry -> new entry or -1), flags and their scans
    int *emap = nullptr, *sflag = nullptr, *spos = nullptr, *cq2 = nullptr;
    size_t emap_cap = 0, sflag_cap = 0, spos_cap = 0, cq2_cap = 0;
    bool with_pairs = true;  // build the covisibility pairs and S blocks (not needed by the matrix-free form)
    bool updated = false;    // the next build() has nothing to do
    int updated_empty = -1;  // first image without entries after that update
    int npts0 = 0, ntr0 = 0;
    // room of the track arrays and their snapshot (upload_tracks reallocates
    // both when a structure needs more), and its host staging
    size_t trk_np_alloc = 0, trk_nt_alloc = 0;
    void *stage = nullptr;  // pinned
    size_t stage_cap = 0;
    // pinned host staging for the structure's small readbacks and uploads
    int *hpin = nullptr;
    size_t hpin_cap = 0;
    int pinned(size_t n_ints);

    int scratch(size_t bytes);
    // working tracks <- snapshot (device-to-device), sizes back to the initial ones
    int restore_tracks(hipStream_t st);
    // tracks of the host structure (after PointXXvIn) plus the gathered
    // observation values and weights
    int upload_tracks(const Structure &s, const float *obs_xy, const std::vector<float> &w0, hipStream_t st);
    // PointXXvIn (rbacore.cpp:1918-1974) on the device from the block's
    // file-order arrays: tracks in file order within a point, points with
    // |Z| <= 1e-8 or fewer than two observations left out; the kept points'
    // coordinates go to X0 (3 * n_pts doubles of room).  Returns
    // RPBA_E_INVALID for an index out of range or a camera twice in one track.
    int build_tracks_device(const rpba_block *io, double *X0, hipStream_t st);
    // camera-major order, pairs and blocks; *first_empty = first image without
    // observations (then nothing else is built) or -1
    int build(hipStream_t st, int *first_empty);
    int chunks_from_camoff(hipStream_t st, int *first_empty);
    int update_after_delete(int ntr_old, hipStream_t st);
    int point_groups(hipStream_t st);
    // RobustDelete on the flags of the stored state; X (accepted points) is
    // compacted in place
    int robust_delete(const unsigned char *flag, int min_keep, double *X, hipStream_t st);
    void swap_tracks(Tracks &b);
    int download_tracks(std::vector<int> &toff_h, std::vector<int> &nrpv_h, std::vector<int> &tfo_h,
                        std::vector<int> &tcam_h, hipStream_t st) const;
    void release();
};

}  // namespace rpba
