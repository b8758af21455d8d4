// Host-side index structures; see structure.h.
#include "structure.h"

#include <algorithm>
#include <cmath>

namespace rpba {

int build_tracks(const rpba_block *io, Structure &s)
{
    const int nc = io->n_cams, np = io->n_pts, no = io->n_obs;
    s.nc = nc;
    std::vector<int> cnt((size_t)np + 1, 0);
    for (int i = 0; i < no; ++i) {
        const int c = io->obs_cam[i], p = io->obs_pt[i];
        if (c < 0) continue;  // IMAPNR.image = -1: not part of the block
        if (c >= nc || p < 0 || p >= np) return RPBA_E_INVALID;
        cnt[(size_t)p + 1]++;
    }
    for (int i = 0; i < np; ++i) cnt[(size_t)i + 1] += cnt[i];
    std::vector<int> all((size_t)cnt[np]);
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int i = 0; i < no; ++i)
            if (io->obs_cam[i] >= 0) all[(size_t)pos[io->obs_pt[i]]++] = i;
    }
    s.nrpv.clear(); s.toff.clear(); s.tcam.clear(); s.tpt.clear(); s.tfo.clear();
    s.toff.push_back(0);
    std::vector<int> seen((size_t)nc, -1);
    for (int i = 0; i < np; ++i) {
        const int t = cnt[(size_t)i + 1] - cnt[i];
        for (int j = cnt[i]; j < cnt[(size_t)i + 1]; ++j) {
            const int c = io->obs_cam[all[j]];
            if (seen[c] == i) return RPBA_E_INVALID;  // duplicate camera in a track
            seen[c] = i;
        }
        if (!(std::fabs((double)io->pts[3 * (size_t)i + 2]) > 1.e-8 && t >= 2)) continue;
        const int kept = (int)s.nrpv.size();
        s.nrpv.push_back(i);
        for (int j = cnt[i]; j < cnt[(size_t)i + 1]; ++j) {
            s.tcam.push_back(io->obs_cam[all[j]]);
            s.tpt.push_back(kept);
            s.tfo.push_back(all[j]);
        }
        s.toff.push_back((int)s.tcam.size());
    }
    s.npts = (int)s.nrpv.size();
    s.ntr = (int)s.tcam.size();
    return RPBA_OK;
}

int build_camera_major(Structure &s)
{
    const int nc = s.nc;
    s.camoff.assign((size_t)nc + 1, 0);
    for (int q = 0; q < s.ntr; ++q) s.camoff[(size_t)s.tcam[q] + 1]++;
    int empty = -1;
    for (int k = 0; k < nc; ++k)
        if (s.camoff[(size_t)k + 1] == 0 && empty < 0) empty = k;
    for (int k = 0; k < nc; ++k) s.camoff[(size_t)k + 1] += s.camoff[k];
    s.camq.assign((size_t)s.ntr, 0);
    std::vector<int> pos(s.camoff.begin(), s.camoff.end() - 1);
    for (int q = 0; q < s.ntr; ++q) s.camq[(size_t)pos[s.tcam[q]]++] = q;
    s.max_track = 0;
    for (int i = 0; i < s.npts; ++i) s.max_track = std::max(s.max_track, s.toff[(size_t)i + 1] - s.toff[i]);
    return empty;
}

// stable counting sort of idx by key[idx] in [0, nkeys)
static void stable_sort_by(std::vector<int> &idx, const std::vector<int> &key, int nkeys)
{
    std::vector<int64_t> cnt((size_t)nkeys + 1, 0);
    for (int i : idx) cnt[(size_t)key[i] + 1]++;
    for (int k = 0; k < nkeys; ++k) cnt[(size_t)k + 1] += cnt[k];
    std::vector<int> out(idx.size());
    for (int i : idx) out[(size_t)cnt[key[i]]++] = i;
    idx.swap(out);
}

void build_pairs(Structure &s)
{
    const int nc = s.nc;
    int64_t P = 0;
    for (int i = 0; i < s.npts; ++i) {
        const int64_t t = s.toff[(size_t)i + 1] - s.toff[i];
        P += t * (t - 1) / 2;
    }
    std::vector<int> a((size_t)P), b((size_t)P), kk((size_t)P), ll((size_t)P);
    int64_t n = 0;
    for (int i = 0; i < s.npts; ++i)
        for (int x = s.toff[i]; x < s.toff[(size_t)i + 1]; ++x)
            for (int y = x + 1; y < s.toff[(size_t)i + 1]; ++y) {
                const bool sw = s.tcam[x] > s.tcam[y];
                a[n] = sw ? y : x;
                b[n] = sw ? x : y;
                kk[n] = s.tcam[a[n]];
                ll[n] = s.tcam[b[n]];
                ++n;
            }
    // LSD radix by (k, l): stable by l, then stable by k
    std::vector<int> idx((size_t)P);
    for (int64_t i = 0; i < P; ++i) idx[i] = (int)i;
    stable_sort_by(idx, ll, nc);
    stable_sort_by(idx, kk, nc);
    s.pa.resize((size_t)P); s.pb.resize((size_t)P);
    s.blk_k.clear(); s.blk_l.clear(); s.blk_poff.clear();
    for (int64_t j = 0; j < P; ++j) {
        const int i = idx[j];
        s.pa[j] = a[i];
        s.pb[j] = b[i];
        if (j == 0 || kk[i] != s.blk_k.back() || ll[i] != s.blk_l.back()) {
            s.blk_k.push_back(kk[i]);
            s.blk_l.push_back(ll[i]);
            s.blk_poff.push_back((int)j);
        }
    }
    s.blk_poff.push_back((int)P);
    const int nb = (int)s.blk_k.size();
    s.row_off.assign((size_t)nc + 1, 0);
    s.col_off.assign((size_t)nc + 1, 0);
    for (int bb = 0; bb < nb; ++bb) {
        s.row_off[(size_t)s.blk_k[bb] + 1]++;
        s.col_off[(size_t)s.blk_l[bb] + 1]++;
    }
    for (int k = 0; k < nc; ++k) {
        s.row_off[(size_t)k + 1] += s.row_off[k];
        s.col_off[(size_t)k + 1] += s.col_off[k];
    }
    s.col_blk.assign((size_t)nb, 0);
    std::vector<int> pos(s.col_off.begin(), s.col_off.end() - 1);
    for (int bb = 0; bb < nb; ++bb) s.col_blk[(size_t)pos[s.blk_l[bb]]++] = bb;  // k ascending within a column
}

void robust_delete(Structure &s, const std::vector<uint8_t> &flag, int min_keep,
                   std::vector<float> &w0, std::vector<double> &X)
{
    std::vector<int> nrpv, toff(1, 0), tcam, tpt, tfo;
    std::vector<float> nw;
    std::vector<double> nX;
    for (int i = 0; i < s.npts; ++i) {
        const int b = s.toff[i], e = s.toff[(size_t)i + 1];
        int nf = 0;
        for (int q = b; q < e; ++q) nf += flag[q] ? 1 : 0;
        if (nf > 0 && (e - b) - nf < min_keep) continue;  // point erased
        const int kept = (int)nrpv.size();
        for (int q = b; q < e; ++q) {
            if (nf > 0 && flag[q]) continue;
            tcam.push_back(s.tcam[q]);
            tpt.push_back(kept);
            tfo.push_back(s.tfo[q]);
            for (int j = 0; j < 3; ++j) nw.push_back(w0[3 * (size_t)q + j]);
        }
        toff.push_back((int)tcam.size());
        nrpv.push_back(s.nrpv[i]);
        for (int j = 0; j < 3; ++j) nX.push_back(X[3 * (size_t)i + j]);
    }
    s.nrpv.swap(nrpv); s.toff.swap(toff); s.tcam.swap(tcam); s.tpt.swap(tpt); s.tfo.swap(tfo);
    w0.swap(nw);
    X.swap(nX);
    s.npts = (int)s.nrpv.size();
    s.ntr = (int)s.tcam.size();
}

}  // namespace rpba
