// Host-side index structures of one bundle block: tracks, camera-major order,
// covisibility pairs grouped by reduced-camera-system block, and the robust
// deletion that rebuilds them.  Pure C++ (no HIP), built once per LM pass.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/rpba.h"

namespace rpba {

struct Structure {
    int nc = 0;      // images (= cameras)
    int npts = 0;    // kept points
    int ntr = 0;     // track entries (observations in kept tracks)
    // point-major tracks: entries of point i are [toff[i], toff[i+1])
    std::vector<int> nrpv;  // kept point -> original point index
    std::vector<int> toff;  // [npts + 1]
    std::vector<int> tcam;  // [ntr] image of the entry
    std::vector<int> tpt;   // [ntr] kept point of the entry
    std::vector<int> tfo;   // [ntr] file-order observation index
    // camera-major order (the reference's nro numbering, rbacore.cpp:745-757):
    // entries of image k are camq[camoff[k] .. camoff[k+1]) in point order
    std::vector<int> camoff;  // [nc + 1]
    std::vector<int> camq;    // [ntr]
    // point pairs (a, b), cam(a) < cam(b), grouped by S block (k, l) in
    // row-major block order and in (point, track position) order inside a block
    std::vector<int> pa, pb;        // [P]
    std::vector<int> blk_k, blk_l;  // [nb] upper off-diagonal blocks, k < l
    std::vector<int> blk_poff;      // [nb + 1] pairs of block b
    std::vector<int> row_off;       // [nc + 1] blocks of row k
    std::vector<int> col_off;       // [nc + 1] blocks of column l ...
    std::vector<int> col_blk;       // [nb]    ... as block ids sorted by k
    int max_track = 0;
    int64_t npairs() const { return (int64_t)pa.size(); }
    int nblocks() const { return (int)blk_k.size(); }
};

// PointXXvIn (rbacore.cpp:1918-1974) on the flat block: tracks in file order,
// points with |Z| <= 1e-8 or fewer than 2 observations dropped.  Returns
// RPBA_E_INVALID on out-of-range indices or a camera seen twice in a track.
int build_tracks(const rpba_block *io, Structure &s);

// camera-major order and per-image counts; returns the first image without
// observations or -1 (rbacore.cpp:401-425)
int build_camera_major(Structure &s);

// covisibility pairs and the block structure of S (rbacore.cpp:1350-1392
// builds the same blocks on the fly in its profile store)
void build_pairs(Structure &s);

// RobustDelete (rbacore.cpp:922-995): drop flagged entries, drop a point whose
// remaining track is shorter than min_keep.  flag is indexed by track entry.
// w0 (3 per entry) and X (3 per point) are compacted alongside.
void robust_delete(Structure &s, const std::vector<uint8_t> &flag, int min_keep,
                   std::vector<float> &w0, std::vector<double> &X);

}  // namespace rpba
