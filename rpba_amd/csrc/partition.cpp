// The partitioned RPBA solve (CalcPLSA's parallel branch, rpbacore.cpp:100-510)
// on one GPU, and the image partition that replaces METIS.
//
// Outer loop (host, as the reference): per outer iteration every part's LSA
// takes one damped Gauss-Newton step in GCP mode (no gauge camera, its
// shared points held by priors built from the other parts' information,
// rbacore.cpp:233-329, 1677-1729), then every point is re-triangulated by
// CalcLSAXMulti (rpbacore.cpp:536-773).  On the device:
//  - the parts' steps run as ONE launch sequence over the union of the parts
//    (cameras of a part contiguous, a part's points duplicated per part, so S
//    is block diagonal): the single-block linearisation, Schur and
//    back-substitution kernels unchanged, the point priors in k_lin_points,
//    and a block-segmented PCG in which every part keeps its own scalars,
//    best iterate and stopping rule (partition_kernels.hip);
//  - the point pass is one thread per point (k_lsax).
// The bookkeeping between them (the tracks regrouped part by part, the
// pitflags, the GCP numbering and the per-part weight lists, the robust
// medians, the best state) runs on the host in the reference's order.
#include "partition.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <vector>

#include "host_util.h"
#include "structure.h"
#include "structure_dev.h"

namespace rpba {

// ---------------------------------------------------------------------------
// the image partition
// ---------------------------------------------------------------------------

namespace {

struct Graph {
    int n = 0, words = 0;
    std::vector<uint64_t> adj;  // n x words bitset: images sharing a point
    std::vector<int> w;         // vertex weights
    bool has(int a, int b) const { return (adj[(size_t)a * words + (b >> 6)] >> (b & 63)) & 1u; }
};

// vertices reachable from s inside the set (mask), breadth first, ties by index
static std::vector<int> bfs_order(const Graph &g, int s, const std::vector<char> &in)
{
    // the set's vertices not yet reached, as bits: a vertex's row masked by
    // them gives exactly its new neighbours, in index order
    std::vector<uint64_t> avail(g.words, 0);
    for (int v = 0; v < g.n; ++v)
        if (in[v]) avail[v >> 6] |= 1ull << (v & 63);
    std::vector<int> order;
    order.push_back(s);
    avail[s >> 6] &= ~(1ull << (s & 63));
    for (size_t h = 0; h < order.size(); ++h) {
        const int v = order[h];
        const uint64_t *row = g.adj.data() + (size_t)v * g.words;
        for (int wd = 0; wd < g.words; ++wd) {
            uint64_t m = row[wd] & avail[wd];
            if (!m) continue;
            avail[wd] &= ~m;
            while (m) {
                const int b = __builtin_ctzll(m);
                m &= m - 1;
                order.push_back(wd * 64 + b);
            }
        }
    }
    return order;
}

// edges of v into the vertices marked side
static int edges_to(const Graph &g, int v, const std::vector<uint64_t> &sidebits)
{
    int c = 0;
    const uint64_t *row = g.adj.data() + (size_t)v * g.words;
    for (int wd = 0; wd < g.words; ++wd) c += __builtin_popcountll(row[wd] & sidebits[wd]);
    return c;
}

// split the vertices vs into parts first .. first + k - 1 by recursive
// bisection: graph growing from a pseudo-peripheral vertex to the weight
// share of the first half, then boundary moves that cut fewer edges while the
// weights stay within the balance
static void bisect(const Graph &g, const std::vector<int> &vs, int k, int first, std::vector<int> &part)
{
    if (k <= 1 || vs.size() <= 1) {
        for (int v : vs) part[v] = first;
        return;
    }
    const int k1 = k / 2, k2 = k - k1;
    long long W = 0;
    for (int v : vs) W += g.w[v];
    const double target = (double)W * (double)k1 / (double)k;
    std::vector<char> in(g.n, 0);
    for (int v : vs) in[v] = 1;
    std::vector<char> side(g.n, 0);  // 1 = first half
    // grow from the vertex farthest (breadth first) from the lowest index;
    // vertices in other components follow in index order
    std::vector<int> order;
    {
        std::vector<char> done(g.n, 0);
        int rem = (int)vs.size();
        for (int v0 : vs) {
            if (done[v0]) continue;
            std::vector<int> o1 = bfs_order(g, v0, in);
            const int far = o1.back();
            std::vector<int> o2 = bfs_order(g, far, in);
            for (int v : o2)
                if (!done[v]) {
                    done[v] = 1;
                    order.push_back(v);
                    --rem;
                }
            if (rem == 0) break;
        }
    }
    long long acc = 0;
    size_t cut = 0;
    const size_t need_right = (size_t)k2, need_left = (size_t)k1;
    for (; cut < order.size(); ++cut) {
        if ((double)acc >= target && cut >= need_left) break;
        if (order.size() - cut <= need_right) break;
        acc += g.w[order[cut]];
    }
    for (size_t j = 0; j < cut; ++j) side[order[j]] = 1;
    // boundary refinement: move a vertex across when it has more edges on the
    // other side and the move keeps both halves within 3% of their targets
    long long wl = acc;
    const double tol = 0.03 * (double)W + 1.;
    for (int pass = 0; pass < 4; ++pass) {
        std::vector<uint64_t> lb(g.words, 0), rb(g.words, 0);
        size_t nl = 0;
        for (int v : vs) {
            if (side[v]) {
                lb[v >> 6] |= 1ull << (v & 63);
                ++nl;
            } else {
                rb[v >> 6] |= 1ull << (v & 63);
            }
        }
        size_t nr = vs.size() - nl;
        bool moved = false;
        for (int v : vs) {
            const int el = edges_to(g, v, lb), er = edges_to(g, v, rb);
            if (side[v] && er > el) {
                if (nl <= need_left || std::fabs((double)(wl - g.w[v]) - target) > tol) continue;
                side[v] = 0;
                wl -= g.w[v];
                lb[v >> 6] &= ~(1ull << (v & 63));
                rb[v >> 6] |= 1ull << (v & 63);
                --nl;
                ++nr;
                moved = true;
            } else if (!side[v] && el > er) {
                if (nr <= need_right || std::fabs((double)(wl + g.w[v]) - target) > tol) continue;
                side[v] = 1;
                wl += g.w[v];
                rb[v >> 6] &= ~(1ull << (v & 63));
                lb[v >> 6] |= 1ull << (v & 63);
                ++nl;
                --nr;
                moved = true;
            }
        }
        if (!moved) break;
    }
    std::vector<int> a, b;
    for (int v : vs) (side[v] ? a : b).push_back(v);
    bisect(g, a, k1, first, part);
    bisect(g, b, k2, first + k1, part);
}

}  // namespace

int partition_images(int nc, int nobs, const int *obs_cam, const int *obs_pt, int npts, int nparts, int *part)
{
    if (nc < 1 || nparts < 1 || nparts > nc || !obs_cam || !obs_pt || !part || nobs < 0 || npts < 0) return RPBA_E_INVALID;
    // tracks (file order), then the covisibility bitsets and the weights
    // BM = sum over the image's points of their track length
    // (PartitionpointXXv, rpbacore.cpp:956-1003)
    std::vector<int> cnt((size_t)npts + 1, 0);
    for (int q = 0; q < nobs; ++q) {
        const int c = obs_cam[q], p = obs_pt[q];
        if (c < 0) continue;
        if (c >= nc || p < 0 || p >= npts) return RPBA_E_INVALID;
        cnt[(size_t)p + 1]++;
    }
    for (int i = 0; i < npts; ++i) cnt[(size_t)i + 1] += cnt[i];
    std::vector<int> cams((size_t)cnt[npts]);
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int q = 0; q < nobs; ++q)
            if (obs_cam[q] >= 0) cams[(size_t)pos[obs_pt[q]]++] = obs_cam[q];
    }
    Graph g;
    g.n = nc;
    g.words = (nc + 63) / 64;
    g.adj.assign((size_t)nc * g.words, 0);
    // bits and sums from ranges of points at once: OR and integer addition
    // give the same result in any order
    std::vector<long long> BM((size_t)nc, 0);
    uint64_t *adj = g.adj.data();
    long long *bm = BM.data();
    const int words = g.words;
    parallel_ranges(npts, host_threads(), [&](long long i0, long long i1, int) {
        for (long long i = i0; i < i1; ++i) {
            const int b = cnt[i], e = cnt[(size_t)i + 1], t = e - b;
            if (t < 2) continue;
            for (int x = b; x < e; ++x) {
                const int fi = cams[x];
                __atomic_fetch_add(bm + fi, (long long)t, __ATOMIC_RELAXED);
                for (int y = x + 1; y < e; ++y) {
                    const int si = cams[y];
                    if (si == fi) continue;
                    __atomic_fetch_or(adj + (size_t)fi * words + (si >> 6), 1ull << (si & 63), __ATOMIC_RELAXED);
                    __atomic_fetch_or(adj + (size_t)si * words + (fi >> 6), 1ull << (fi & 63), __ATOMIC_RELAXED);
                }
            }
        }
    });
    g.w.resize(nc);
    for (int i = 0; i < nc; ++i) g.w[i] = std::max(1, (int)std::pow((double)BM[i], 0.3333));
    std::vector<int> vs(nc), pv(nc, 0);
    std::iota(vs.begin(), vs.end(), 0);
    bisect(g, vs, nparts, 0, pv);
    for (int i = 0; i < nc; ++i) part[i] = pv[i];
    return RPBA_OK;
}

// ---------------------------------------------------------------------------
// the partitioned solve
// ---------------------------------------------------------------------------

namespace {

// RPBA_PART_TRACE=1: the phases of every outer iteration with their wall
// times on stderr (the stream synchronised at each mark)
struct PhaseClock {
    bool on = false;
    std::chrono::steady_clock::time_point t;
    PhaseClock()
    {
        const char *e = std::getenv("RPBA_PART_TRACE");
        on = e && e[0] == '1';
        t = std::chrono::steady_clock::now();
    }
    void mark(hipStream_t st, const char *what, int iter)
    {
        if (!on) return;
        if (st) (void)hipStreamSynchronize(st);
        const auto n = std::chrono::steady_clock::now();
        std::fprintf(stderr, "rpba part: iter %2d %-26s %9.1f ms\n", iter, what, std::chrono::duration<double, std::milli>(n - t).count());
        t = n;
    }
};

struct PartSolver {
    rpba_block *io = nullptr;
    rpba_log_fn log = nullptr;
    void *user = nullptr;
    int N = 0, pxs = 0, no = 0, P = 0, robust = 0, addf0[3] = {0, 0, 0}, nraddp = 0;
    float ihw2 = 1.f;
    double defdiff = 5000.;
    hipStream_t st = nullptr;
    std::vector<int> part, pcnt, pimg;
    // tracks: point i's entries trk[base[i] .. base[i] + len[i])
    std::vector<long long> base, soff;
    std::vector<int> len, trk, spart, gcpinfo;
    std::vector<unsigned char> pit, bkept;
    std::vector<double> XXd, PMd, gcpc;
    std::vector<float> cf, cd, sig0wv;
    std::vector<std::vector<float>> w3list;  // per part, 6 per GCP entry (interleaved)
    std::vector<float> w3h;                  // host copy of the slots' information
    // device: static inputs of the point pass and its state
    DBuf<long long> d_base, d_soff;
    DBuf<int> d_len, d_trk, d_obs_cam, d_part, d_spart, d_nout, d_rd;
    DBuf<double> d_X, d_PMd, d_nv, d_sg;
    DBuf<float> d_cf, d_cd, d_oxy, d_ow, d_sig0wv, d_w3, d_wk, d_keys;
    DBuf<unsigned char> d_pit, d_dl;
    int64_t pcg_total = 0;
    PhaseClock clk;
    int cur_iter = 0;
    int nthr = host_threads();  // host threads of the bookkeeping loops (per-point ranges)

    void logs(const char *s)
    {
        if (log) log(s, user);
    }
    std::vector<int> ecam, eslot;  // per entry: image, slot (entry_index)
    // the union of the parts and its device buffers, kept from one outer
    // iteration to the next (their room only grows)
    std::unique_ptr<DevStruct> uds;
    struct UnionBufs {
        DBuf<double> Wrec, Yrec, Vinv, gp, Sd, Su, bv, Dd, rowpart, colpart, dpart, Cinv, x, best, r, dv, Ad, h, h1, X, Xn,
            Sdf, ev, c_Rb, c_Cb, c_Rn, c_Cn, c_In, c_th, c_thn, c_Pst, c_Ist, gc_c;
        DBuf<float> gc_w;
        DBuf<unsigned> c_mask;
        DBuf<unsigned char> camflag;
        DBuf<int> gc_of, d_boff, d_cblk, d_ptblk, d_nan, d_ndone, d_live;
        DBuf<SegScal> segS;
        DBuf<PcgScalars> pcgS;
        void release()
        {
            for (DBuf<double> *b : {&Wrec, &Yrec, &Vinv, &gp, &Sd, &Su, &bv, &Dd, &rowpart, &colpart, &dpart, &Cinv, &x, &best, &r,
                                    &dv, &Ad, &h, &h1, &X, &Xn, &Sdf, &ev, &c_Rb, &c_Cb, &c_Rn, &c_Cn, &c_In, &c_th, &c_thn,
                                    &c_Pst, &c_Ist, &gc_c})
                b->release();
            gc_w.release();
            c_mask.release();
            camflag.release();
            for (DBuf<int> *b : {&gc_of, &d_boff, &d_cblk, &d_ptblk, &d_nan, &d_ndone, &d_live}) b->release();
            segS.release();
            pcgS.release();
        }
    } ub;
    Structure us;                                 // the union's host structure
    std::vector<int> up, ptblk, gcp_of, scnt;     // slot -> union point; union point -> block, prior; entries per slot
    std::vector<double> Xu, gcc;                  // union points' X; the priors' centres
    std::vector<float> w0, gcw;                   // union entries' weights; the priors' weights
    ~PartSolver()
    {
        ub.release();
        if (uds) uds->release();
    }
    int point_pass(int robustmode, int weightonly, int wxyflag, std::vector<double> &sg, std::vector<int> &rd,
                   std::vector<float> *keys);
    int block_step(int iter, const int addf[3], int &pcg_it);
    void entry_index();
};

// CalcLSAXMulti over every point with a track (rpbacore.cpp:311-392)
int PartSolver::point_pass(int robustmode, int weightonly, int wxyflag, std::vector<double> &sg, std::vector<int> &rd,
                           std::vector<float> *keys)
{
    const size_t ntr = (size_t)base[pxs];
    int rc = 0;
    rc |= d_len.put(len.data(), len.size(), st);
    rc |= d_trk.put(trk.data(), std::max<size_t>(ntr, 1), st);
    rc |= d_X.put(XXd.data(), XXd.size(), st);
    rc |= d_PMd.put(PMd.data(), PMd.size(), st);
    rc |= d_cf.put(cf.data(), cf.size(), st);
    rc |= d_cd.put(cd.data(), cd.size(), st);
    rc |= d_pit.put(pit.data(), std::max<size_t>(pit.size(), 1), st);
    rc |= d_sig0wv.put(sig0wv.data(), sig0wv.size(), st);
    if (rc) return RPBA_E_NOMEM;
    LsaxArgs a;
    a.npts = pxs; a.P = P; a.robustmode = robustmode; a.weightonly = weightonly; a.wxyflag = wxyflag;
    a.base = d_base.p; a.len = d_len.p; a.trk = d_trk.p; a.X = d_X.p; a.PMd = d_PMd.p; a.cf = d_cf.p; a.cd = d_cd.p;
    a.oxy = d_oxy.p; a.ow = io->obs_w ? d_ow.p : nullptr; a.obs_cam = d_obs_cam.p; a.part = d_part.p; a.sig0wv = d_sig0wv.p;
    a.soff = d_soff.p; a.spart = d_spart.p; a.pit = d_pit.p; a.w3 = d_w3.p; a.nv = d_nv.p; a.wk = d_wk.p; a.dl = d_dl.p;
    a.keys = keys ? d_keys.p : nullptr; a.nout = d_nout.p; a.sg = d_sg.p; a.rd = d_rd.p;
    launch_lsax(a, st);
    HIPCHK(hipGetLastError());
    std::vector<int> nout((size_t)pxs);
    sg.resize((size_t)pxs);
    rd.resize((size_t)pxs);
    if (pxs) {
        HIPCHK(hipMemcpyAsync(nout.data(), d_nout.p, sizeof(int) * (size_t)pxs, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(sg.data(), d_sg.p, sizeof(double) * (size_t)pxs, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(rd.data(), d_rd.p, sizeof(int) * (size_t)pxs, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(XXd.data(), d_X.p, sizeof(double) * XXd.size(), hipMemcpyDeviceToHost, st));
    }
    if (robustmode == 2 && ntr) HIPCHK(hipMemcpyAsync(trk.data(), d_trk.p, sizeof(int) * ntr, hipMemcpyDeviceToHost, st));
    if (keys) {
        keys->resize(std::max<size_t>(ntr, 1));
        if (ntr) HIPCHK(hipMemcpyAsync(keys->data(), d_keys.p, sizeof(float) * ntr, hipMemcpyDeviceToHost, st));
    }
    if (wxyflag && !w3h.empty())
        HIPCHK(hipMemcpyAsync(w3h.data(), d_w3.p, sizeof(float) * w3h.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < pxs; ++i) len[i] = nout[i];
    return RPBA_OK;
}

// every valid part's LSA step (GCP mode, one iteration, rbacore.cpp:333-616 with
// itbreakflag; OutputLSAP 620-640) as one launch sequence over their union
// per current entry e (trk positions base[i] .. base[i] + len[i]): its
// image and its slot, the point's slot of the entry's part
void PartSolver::entry_index()
{
    ecam.resize(trk.size());
    eslot.resize(trk.size());
    parallel_ranges(pxs, nthr, [&](long long i0, long long i1, int) {
        for (long long i = i0; i < i1; ++i) {
            const long long b = base[i], s0 = soff[i];
            for (int q = 0; q < len[i]; ++q) {
                const int c = io->obs_cam[trk[(size_t)(b + q)]];
                ecam[(size_t)(b + q)] = c;
                const int p = part[c];
                long long s = s0;
                while (spart[(size_t)s] != p) ++s;  // the slots hold every part of the initial track
                eslot[(size_t)(b + q)] = (int)s;
            }
        }
    });
}

int PartSolver::block_step(int iter, const int addf[3], int &pcg_it)
{
    pcg_it = 0;
    const int nraddpi = addf[0] + addf[1] + addf[2], nrpari = 6 + nraddpi;
    const size_t NS = spart.size();
    entry_index();
    // entries of each point in each of its parts
    // (per-point loops over ranges of points: a point's slots and entries are
    // its own; per-part and per-image counts are summed over the ranges)
    const int nr = nthr;
    scnt.assign(NS, 0);
    parallel_ranges(pxs, nr, [&](long long i0, long long i1, int) {
        for (long long i = i0; i < i1; ++i)
            for (int q = 0; q < len[i]; ++q) scnt[(size_t)eslot[(size_t)(base[i] + q)]]++;
    });
    // PointXXvIn with GCP information (rbacore.cpp:1949-1962): a point stays
    // in a part with two entries there, or one if its GCP number is > 0;
    // validity of each part (an image without entries, or a negative
    // redundancy, ends its Adjust before the step: rbacore.cpp:401-430, 466-476)
    std::vector<char> valid((size_t)P, 1), lowred((size_t)P, 0);
    std::vector<int> npk((size_t)P, 0), ntk((size_t)P, 0), ngk((size_t)P, 0);
    std::vector<int> rcnt((size_t)nr * (3 * (size_t)P + (size_t)N), 0);  // per range: npk, ntk, ngk, imgcnt
    parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
        int *rp = rcnt.data() + (size_t)r * (3 * (size_t)P + (size_t)N);
        int *rimg = rp + 3 * (size_t)P;
        for (long long i = i0; i < i1; ++i) {
            if (!len[i]) continue;
            const bool zok = std::fabs(XXd[3 * (size_t)i + 2]) > 1.e-8;
            for (long long sl = soff[i]; sl < soff[(size_t)i + 1]; ++sl) {
                const int c = scnt[(size_t)sl];
                if (!c) continue;
                const bool keep = zok && (c >= 2 || (gcpinfo[i] > 0 && c > 0));
                bkept[(size_t)sl] = keep ? 1 : 0;
                if (!keep) continue;
                const int p = spart[(size_t)sl];
                rp[p]++;
                rp[P + p] += c;
                if (gcpinfo[i] > -1) rp[2 * P + p]++;
            }
            for (int q = 0; q < len[i]; ++q) {
                const size_t e = (size_t)(base[i] + q);
                if (bkept[(size_t)eslot[e]]) rimg[ecam[e]]++;
            }
        }
    });
    std::vector<int> imgcnt((size_t)N, 0);
    for (int r = 0; r < nr; ++r) {
        const int *rp = rcnt.data() + (size_t)r * (3 * (size_t)P + (size_t)N);
        for (int p = 0; p < P; ++p) {
            npk[p] += rp[p];
            ntk[p] += rp[P + p];
            ngk[p] += rp[2 * P + p];
        }
        for (int k = 0; k < N; ++k) imgcnt[k] += rp[3 * P + k];
    }
    for (int p = 0; p < P; ++p) {
        for (int k = pcnt[p]; k < pcnt[(size_t)p + 1]; ++k)
            if (imgcnt[pimg[k]] == 0) valid[p] = 0;
        if (valid[p]) {
            const long long nrobs = 2LL * ntk[p] + 3LL * ngk[p];
            const long long nrunk = (long long)nrpari * (pcnt[(size_t)p + 1] - pcnt[p]) + 3LL * npk[p] + (long long)nraddpi * N;
            if (nrobs - nrunk < 0) {
                valid[p] = 0;
                lowred[p] = 1;
            }
        }
    }
    // PrepareIteration re-orthonormalises image 0 of a part on the constructor
    // path (the second outer iteration); EncodeCameras then every image of a
    // part it reaches (rbacore.cpp:151-152, 1025-1031)
    for (int p = 0; p < P; ++p) {
        if (iter == 2) rectify34(PMd.data() + 12 * (size_t)pimg[pcnt[p]]);
        if (valid[p] || lowred[p])
            for (int k = pcnt[p]; k < pcnt[(size_t)p + 1]; ++k) rectify34(PMd.data() + 12 * (size_t)pimg[k]);
    }
    // the union of the valid parts: cameras part by part, points part-major
    // in global order, entries in track order
    std::vector<int> ucam, boff(1, 0), ublk_of;  // union camera -> image; block -> first union camera
    std::vector<int> bpart, pblk((size_t)P, -1);  // union block -> part; part -> union block
    std::vector<int> uidx((size_t)N, -1);
    std::vector<long long> uoff((size_t)P + 1, 0);
    for (int p = 0; p < P; ++p) {
        uoff[(size_t)p + 1] = uoff[p] + (valid[p] ? npk[p] : 0);
        if (!valid[p]) continue;
        pblk[p] = (int)bpart.size();
        for (int k = pcnt[p]; k < pcnt[(size_t)p + 1]; ++k) {
            uidx[pimg[k]] = (int)ucam.size();
            ucam.push_back(pimg[k]);
            ublk_of.push_back((int)bpart.size());
        }
        bpart.push_back(p);
        boff.push_back((int)ucam.size());
    }
    const int nb = (int)bpart.size(), nc = (int)ucam.size();
    clk.mark(st, "block lists and validity", cur_iter);
    if (nb == 0) return RPBA_OK;
    const int nup = (int)uoff[P];
    up.assign(NS, -1);
    us.nc = nc;
    us.npts = nup;
    // every union point's nrpv, toff, Xu and ptblk are written below
    // (resize, no fill); gcp_of takes -1 below for the points without a prior
    us.nrpv.resize((size_t)nup);
    us.toff.resize((size_t)nup + 1);
    us.toff[0] = 0;
    Xu.resize(3 * (size_t)std::max(nup, 1));
    ptblk.resize((size_t)std::max(nup, 1));
    gcp_of.resize((size_t)std::max(nup, 1));
    {
        // union point u of part p: the points of p in global order, so range
        // r of the points starts at p's offset plus the ranges before it
        std::vector<long long> w((size_t)nr * P, 0);
        auto take = [&](long long sl) {
            const int p = spart[(size_t)sl];
            return scnt[(size_t)sl] && bkept[(size_t)sl] && valid[p];
        };
        parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
            long long *wr = w.data() + (size_t)r * P;
            for (long long i = i0; i < i1; ++i)
                if (len[i])
                    for (long long sl = soff[i]; sl < soff[(size_t)i + 1]; ++sl)
                        if (take(sl)) wr[spart[(size_t)sl]]++;
        });
        for (int p = 0; p < P; ++p) {
            long long at = uoff[p];
            for (int r = 0; r < nr; ++r) {
                const long long n = w[(size_t)r * P + p];
                w[(size_t)r * P + p] = at;
                at += n;
            }
        }
        parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
            long long *wr = w.data() + (size_t)r * P;
            for (long long i = i0; i < i1; ++i) {
                if (!len[i]) continue;
                for (long long sl = soff[i]; sl < soff[(size_t)i + 1]; ++sl) {
                    if (!take(sl)) continue;
                    const int p = spart[(size_t)sl];
                    const int u = (int)wr[p]++;
                    up[(size_t)sl] = u;
                    us.nrpv[u] = (int)i;
                    us.toff[(size_t)u + 1] = scnt[(size_t)sl];
                    for (int j = 0; j < 3; ++j) Xu[3 * (size_t)u + j] = XXd[3 * (size_t)i + j];
                    ptblk[u] = pblk[p];
                }
            }
        });
    }
    for (int u = 0; u < nup; ++u) us.toff[(size_t)u + 1] += us.toff[u];
    us.ntr = us.toff[nup];
    clk.mark(st, "union points", cur_iter);
    us.tcam.resize((size_t)us.ntr);  // every union entry is written below
    us.tpt.resize((size_t)us.ntr);
    us.tfo.resize((size_t)us.ntr);
    w0.resize(3 * (size_t)us.ntr);
    {
        // the entries of union point u all come from one point's track
        std::vector<int> pos(us.toff.begin(), us.toff.end() - 1);
        parallel_ranges(pxs, nr, [&](long long i0, long long i1, int) {
            for (long long i = i0; i < i1; ++i)
                for (int q = 0; q < len[i]; ++q) {
                    const size_t e = (size_t)(base[i] + q);
                    const int u = up[(size_t)eslot[e]];
                    if (u < 0) continue;
                    const int at = pos[u]++, fo = trk[e];
                    us.tcam[at] = uidx[ecam[e]];
                    us.tpt[at] = u;
                    us.tfo[at] = fo;
                    w0[3 * (size_t)at] = io->obs_w ? io->obs_w[3 * (size_t)fo] : 1.f;
                    w0[3 * (size_t)at + 1] = io->obs_w ? io->obs_w[3 * (size_t)fo + 1] : 1.f;
                    w0[3 * (size_t)at + 2] = io->obs_w ? io->obs_w[3 * (size_t)fo + 2] : 0.f;
                }
        });
    }
    clk.mark(st, "union entries", cur_iter);
    // the priors: the gn-th kept GCP point of a part takes entry gn of its list
    {
        // the GCP union points in union order: gcp_of[u] = their running
        // count; within a part the count from the part's first union point
        std::vector<long long> rc0((size_t)nr + 1, 0);
        parallel_ranges(nup, nr, [&](long long u0, long long u1, int r) {
            long long n = 0;
            for (long long u = u0; u < u1; ++u) n += gcpinfo[us.nrpv[u]] >= 0;
            rc0[(size_t)r + 1] = n;
        });
        for (int r = 0; r < nr; ++r) rc0[(size_t)r + 1] += rc0[r];
        const long long ng = rc0[nr];
        gcc.resize(3 * (size_t)ng);
        gcw.resize(6 * (size_t)ng);
        parallel_ranges(nup, nr, [&](long long u0, long long u1, int r) {
            long long x = rc0[r];
            for (long long u = u0; u < u1; ++u) {
                const int i = us.nrpv[u];
                if (gcpinfo[i] < 0) {
                    gcp_of[u] = -1;
                    continue;
                }
                gcp_of[u] = (int)x;
                ++x;
            }
        });
        // a part's first GCP union point: its list entry 0
        std::vector<long long> pfirst((size_t)nb, -1);
        for (int bb = 0; bb < nb; ++bb) {
            const int p = bpart[bb];
            for (long long u = uoff[p]; u < uoff[(size_t)p + 1]; ++u)
                if (gcp_of[u] >= 0) {
                    pfirst[bb] = gcp_of[u];
                    break;
                }
        }
        parallel_ranges(nup, nr, [&](long long u0, long long u1, int) {
            for (long long u = u0; u < u1; ++u) {
                const int x = gcp_of[u];
                if (x < 0) continue;
                const int i = us.nrpv[u], bb = ptblk[u];
                const std::vector<float> &L = w3list[(size_t)bpart[bb]];
                const size_t gn = (size_t)(x - pfirst[bb]);
                for (int j = 0; j < 3; ++j) gcc[3 * (size_t)x + j] = gcpc[3 * (size_t)gcpinfo[i] + j];
                for (int v = 0; v < 6; ++v) gcw[6 * (size_t)x + v] = 6 * gn + v < L.size() ? L[6 * gn + v] : 0.f;
            }
        });
    }
    // camera state of the pass (EncodeCameras in GCP mode: every camera has its
    // six pose slots and the estimated intrinsics; no gauge camera)
    std::vector<double> Rb(9 * (size_t)nc), Cb(3 * (size_t)nc), th(9 * (size_t)nc, 0.), In(3 * (size_t)nc),
        Pst(12 * (size_t)nc);
    std::vector<unsigned> mask((size_t)nc);
    for (int u = 0; u < nc; ++u) {
        const double *Pm = PMd.data() + 12 * (size_t)ucam[u];
        std::memcpy(Pst.data() + 12 * (size_t)u, Pm, sizeof(double) * 12);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Rb[9 * (size_t)u + r * 3 + c] = Pm[r * 4 + c];
        const double *R = Rb.data() + 9 * (size_t)u;
        for (int c = 0; c < 3; ++c) Cb[3 * (size_t)u + c] = R[0 * 3 + c] * Pm[3] + R[1 * 3 + c] * Pm[7] + R[2 * 3 + c] * Pm[11];
        unsigned m = 0x3fu;
        if (addf[0]) m |= 1u << 6;
        if (addf[1]) m |= 1u << 7;
        if (addf[2]) m |= 1u << 8;
        mask[u] = m;
        In[3 * (size_t)u] = (double)cd[2 * (size_t)ucam[u]];
        In[3 * (size_t)u + 1] = (double)cd[2 * (size_t)ucam[u] + 1];
        In[3 * (size_t)u + 2] = (double)cf[ucam[u]];
        for (int j = 0; j < 3; ++j) th[9 * (size_t)u + 6 + j] = In[3 * (size_t)u + j];
    }
    clk.mark(st, "priors and cameras", cur_iter);
    // device structure of the union (pairs within parts only)
    if (!uds) uds.reset(new DevStruct());
    DevStruct *ds = uds.get();
    ds->with_pairs = true;
    ds->spmv_waves = 4 * spmv_grid(0);
    int rc = ds->upload_tracks(us, io->obs_xy, w0, st);
    if (rc) return rc;
    int empty = -1;
    rc = ds->build(st, &empty);
    if (rc) return rc;
    if (empty >= 0) return RPBA_E_INVALID;  // ruled out above
    if (ds->nslab <= 0 || ds->nchunk <= 0) return RPBA_E_INVALID;  // a part's S without off-diagonal blocks
    clk.mark(st, "union structure (device)", cur_iter);
    const size_t ntr = (size_t)ds->ntr, np = (size_t)ds->npts;
    UnionBufs &U = ub;
    auto &Wrec = U.Wrec, &Yrec = U.Yrec, &Vinv = U.Vinv, &gp = U.gp, &Sd = U.Sd, &Su = U.Su, &bv = U.bv, &Dd = U.Dd;
    auto &rowpart = U.rowpart, &colpart = U.colpart, &dpart = U.dpart, &Cinv = U.Cinv, &x = U.x, &best = U.best, &r = U.r;
    auto &dv = U.dv, &Ad = U.Ad, &h = U.h, &h1 = U.h1, &X = U.X, &Xn = U.Xn, &Sdf = U.Sdf, &ev = U.ev;
    auto &c_Rb = U.c_Rb, &c_Cb = U.c_Cb, &c_Rn = U.c_Rn, &c_Cn = U.c_Cn, &c_In = U.c_In, &c_th = U.c_th, &c_thn = U.c_thn;
    auto &c_Pst = U.c_Pst, &c_Ist = U.c_Ist, &gc_c = U.gc_c;
    auto &gc_w = U.gc_w;
    auto &c_mask = U.c_mask;
    auto &camflag = U.camflag;
    auto &gc_of = U.gc_of, &d_boff = U.d_boff, &d_cblk = U.d_cblk, &d_ptblk = U.d_ptblk, &d_nan = U.d_nan, &d_ndone = U.d_ndone;
    auto &segS = U.segS;
    auto &pcgS = U.pcgS;
    rc = 0;
    rc |= Wrec.alloc((size_t)kRec * std::max<size_t>(ntr, 1));
    rc |= Yrec.alloc((size_t)kRec * std::max<size_t>(ntr, 1));
    rc |= Vinv.alloc(9 * std::max<size_t>(np, 1));
    rc |= gp.alloc(3 * std::max<size_t>(np, 1));
    rc |= Sd.alloc(81 * (size_t)nc);
    rc |= Su.alloc(81 * (size_t)std::max(ds->nblk, 1));
    rc |= bv.alloc(9 * (size_t)nc);
    rc |= Dd.alloc(90 * (size_t)std::max(ds->nch, 1));
    rc |= rowpart.alloc(63 * (size_t)std::max(ds->nchunk, 1));
    rc |= colpart.alloc(9 * (size_t)std::max(ds->nslab, 1));
    rc |= dpart.alloc((size_t)std::max(spmv_grid(ds->nblk), 1));
    for (DBuf<double> *v : {&Cinv}) rc |= v->alloc(81 * (size_t)nc);
    for (DBuf<double> *v : {&x, &best, &r, &dv, &Ad, &h, &h1, &Sdf}) rc |= v->alloc(9 * (size_t)nc);
    rc |= ev.alloc((size_t)nc);
    rc |= X.put(Xu.data(), std::max<size_t>(Xu.size(), 1), st);
    rc |= Xn.alloc(std::max<size_t>(Xu.size(), 1));
    rc |= c_Rb.put(Rb.data(), Rb.size(), st);
    rc |= c_Cb.put(Cb.data(), Cb.size(), st);
    rc |= c_Rn.alloc(9 * (size_t)nc);
    rc |= c_Cn.alloc(3 * (size_t)nc);
    rc |= c_In.put(In.data(), In.size(), st);
    rc |= c_th.put(th.data(), th.size(), st);
    rc |= c_thn.put(th.data(), th.size(), st);
    rc |= c_Pst.put(Pst.data(), Pst.size(), st);
    rc |= c_Ist.put(In.data(), In.size(), st);
    rc |= c_mask.put(mask.data(), mask.size(), st);
    std::vector<unsigned char> cflag((size_t)nc, 1);  // every camera takes the step; none is held fixed
    rc |= camflag.put(cflag.data(), cflag.size(), st);
    rc |= gc_of.put(gcp_of.data(), std::max<size_t>(gcp_of.size(), 1), st);
    rc |= gc_c.put(gcc.data(), std::max<size_t>(gcc.size(), 1), st);
    rc |= gc_w.put(gcw.data(), std::max<size_t>(gcw.size(), 1), st);
    rc |= d_boff.put(boff.data(), boff.size(), st);
    rc |= d_cblk.put(ublk_of.data(), ublk_of.size(), st);
    rc |= d_ptblk.put(ptblk.data(), std::max<size_t>(ptblk.size(), 1), st);
    rc |= d_nan.alloc((size_t)nb);
    rc |= d_ndone.alloc(1);
    std::vector<SegScal> hs((size_t)nb);
    for (int bb = 0; bb < nb; ++bb) {
        std::memset(&hs[bb], 0, sizeof(SegScal));
        hs[bb].msize = std::max(1000, nrpari * (boff[(size_t)bb + 1] - boff[bb]));  // rbacore.cpp:1848-1849
    }
    rc |= segS.put(hs.data(), hs.size(), st);
    PcgScalars z2[2];
    std::memset(z2, 0, sizeof z2);
    rc |= pcgS.put(z2, 2, st);
    if (rc) return RPBA_E_NOMEM;
    HIPCHK(hipMemsetAsync(d_ndone.p, 0, sizeof(int), st));
    CamPtrs cp;
    cp.Rb = c_Rb.p; cp.Cb = c_Cb.p; cp.Rn = c_Rn.p; cp.Cn = c_Cn.p; cp.In = c_In.p; cp.th = c_th.p; cp.thn = c_thn.p;
    cp.Pst = c_Pst.p; cp.Ist = c_Ist.p; cp.mask = c_mask.p;
    // DecodeCameras(XPMatrices): the trial cameras of the linearisation
    StepArgs s0;
    s0.nc = nc; s0.npts = (int)np; s0.npgrp = ds->npgrp; s0.cam = cp; s0.toff = ds->toff; s0.tcam = ds->tcam;
    s0.pgrp = ds->pgrp; s0.pg4 = ds->pg4; s0.Wrec = Wrec.p; s0.rec32 = 0; s0.Vinv = Vinv.p; s0.gp = gp.p; s0.sol = nullptr; s0.X = X.p;
    s0.Xn = Xn.p; s0.S = pcgS.p; s0.camflag = camflag.p;
    launch_decode(s0, st);
    // CalcBundleSep with multfact 1e-4 (rbacore.cpp:371-372, 1094-1195)
    LinArgs la;
    la.nc = nc; la.npts = (int)np; la.ntr = (int)ntr; la.npgrp = ds->npgrp; la.pgrp = ds->pgrp; la.pg4 = ds->pg4;
    la.defdiff = defdiff; la.mult1 = 1. + 1.e-4; la.cam = cp;
    la.toff = ds->toff; la.tcam = ds->tcam; la.tpt = ds->tpt; la.camoff = ds->camoff; la.camq = ds->camq;
    la.ox = ds->ox; la.oy = ds->oy; la.wx = ds->w0x; la.wy = ds->w0y; la.wxy = ds->w0xy; la.X = X.p;
    la.Wrec = Wrec.p; la.Yrec = Yrec.p; la.rec32 = 0; la.Vinv = Vinv.p; la.gp = gp.p; la.Sd = Sd.p; la.bvec = bv.p;
    la.Dd = Dd.p; la.nch = ds->nch; la.ch_start = ds->ch_start; la.ch_off = ds->ch_off; la.Dk = nullptr; la.yrank = nullptr;
    la.gcp_of = gc_of.p; la.gcp_c = gc_c.p; la.gcp_w = gc_w.p;
    SchurArgs sa;
    sa.nblk = ds->nblk; sa.poff = ds->poff; sa.pab = ds->pab; sa.Yrec = Yrec.p; sa.Wrec = Wrec.p; sa.Su = Su.p;
    sa.sm = ds->sumap(); sa.sperm = ds->sperm; sa.nslab = ds->nslab; sa.slab_bs = ds->slab_bs; sa.npairs = ds->npairs;
    sa.rec32 = 0;
    launch_lin_points(la, st);
    launch_schur_pairs(sa, st);
    launch_schur_diag(la, st);
    launch_lin_cameras(la, st);
    launch_bj_inverse(nc, Sd.p, Cinv.p, st);
    HIPCHK(hipGetLastError());
    clk.mark(st, "linearise and Schur", cur_iter);
    // the parts' PCGs (SparsePCG with begini = 0, rbacore.cpp:1807-1914)
    PcgArgs pa;
    std::memset(&pa, 0, sizeof pa);
    pa.nc = nc; pa.msize = 1000; pa.Sd = Sd.p; pa.Su = Su.p; pa.Cinv = Cinv.p; pa.b = bv.p;
    pa.row_off = ds->roff; pa.col_off = ds->coff; pa.col_blk = ds->cblk; pa.blk_k = ds->bk; pa.blk_l = ds->bl;
    pa.x = x.p; pa.best = best.p; pa.r = r.p; pa.d = dv.p; pa.Ad = Ad.p; pa.h = h.p; pa.h1 = h1.p; pa.part = nullptr;
    pa.sm = ds->sumap(); pa.nslab = ds->nslab; pa.nchunk = ds->nchunk; pa.nstrip = ds->nstrip; pa.slab_chunk = ds->slab_chunk;
    pa.slab_bs = ds->slab_bs; pa.slab_l = ds->slab_l; pa.slab_mask = ds->slab_mask; pa.strip_s0 = ds->strip_s0;
    pa.chunk_off = ds->chunk_off; pa.chunk_strip = ds->chunk_strip; pa.cs_off = ds->cs_off; pa.cs_list = ds->cs_list;
    pa.rowpart = rowpart.p; pa.colpart = colpart.p; pa.dpart = dpart.p; pa.nblk = ds->nblk;
    pa.ndpart = spmv_grid(ds->nblk); pa.dsum = dpart.p; pa.ndsum = pa.ndpart; pa.S = pcgS.p; pa.j = 0;
    pa.implicit = 0; pa.npts = (int)np; pa.ntr = (int)ntr; pa.nch = ds->nch;
    SegArgs sg;
    sg.nc = nc; sg.nb = nb; sg.boff = d_boff.p; sg.cblk = d_cblk.p; sg.S = segS.p; sg.ndone = d_ndone.p; sg.Sdf = Sdf.p;
    sg.e = ev.p;
    rc = U.d_live.alloc((size_t)std::max(ds->nstrip, 1));
    if (rc) return RPBA_E_NOMEM;
    sg.strip_live = U.d_live.p;
    sg.nstrip = ds->nstrip;
    pa.strip_live = U.d_live.p;
    launch_seg_init(pa, sg, st);
    int *hd = nullptr;
    if (hipHostMalloc((void **)&hd, sizeof(int), hipHostMallocDefault) != hipSuccess) return RPBA_E_NOMEM;
    int launched = 0, maxits = 0;
    for (const SegScal &q : hs) maxits = std::max(maxits, q.msize);
    rc = RPBA_OK;
    for (;;) {
        for (int j = 0; j < 4; ++j) launch_seg_iter(pa, sg, st);
        launched += 4;
        if (hipMemcpyAsync(hd, d_ndone.p, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = RPBA_E_HIP;
            break;
        }
        if (*hd >= nb) break;
        if (launched > maxits + 8) {  // cannot happen: every block stops at its msize
            rc = RPBA_E_HIP;
            break;
        }
    }
    (void)hipHostFree(hd);
    if (rc) return rc;
    clk.mark(st, "segmented PCG", cur_iter);
    // back-substitution, the per-part NaN check of the step, decode, store
    StepArgs s1 = s0;
    s1.sol = best.p;
    launch_backsub(s1, st);
    launch_seg_nan(nc, (int)np, d_cblk.p, d_ptblk.p, best.p, Xn.p, nb, d_nan.p, st);
    launch_decode(s1, st);
    launch_store_cams(cp, nc, st);
    HIPCHK(hipGetLastError());
    std::vector<int> nanf((size_t)nb);
    HIPCHK(hipMemcpyAsync(Pst.data(), c_Pst.p, sizeof(double) * Pst.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(In.data(), c_Ist.p, sizeof(double) * In.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(nanf.data(), d_nan.p, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hs.data(), segS.p, sizeof(SegScal) * (size_t)nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    clk.mark(st, "step, decode, store", cur_iter);
    // OutputLSAP: the stored matrices and camerasnew (float) of every part
    // that stored; a part whose step was NaN keeps its encoded matrices and
    // its intrinsics (rbacore.cpp:541-546 returns before the decode)
    for (int bb = 0; bb < nb; ++bb) {
        pcg_it += hs[bb].iters;
        if (nanf[bb]) continue;
        for (int u = boff[bb]; u < boff[(size_t)bb + 1]; ++u) {
            const int g = ucam[u];
            std::memcpy(PMd.data() + 12 * (size_t)g, Pst.data() + 12 * (size_t)u, sizeof(double) * 12);
            cd[2 * (size_t)g] = (float)In[3 * (size_t)u];
            cd[2 * (size_t)g + 1] = (float)In[3 * (size_t)u + 1];
            cf[g] = (float)In[3 * (size_t)u + 2];
        }
    }
    return RPBA_OK;
}

}  // namespace

int partitioned_solve(rpba_block *io, const rpba_flags *f, int P, rpba_log_fn log, void *user, rpba_stats *stats)
{
    PartSolver S;
    S.io = io; S.log = log; S.user = user;
    S.N = io->n_cams; S.pxs = io->n_pts; S.no = io->n_obs; S.P = P;
    S.robust = f->robust != 0;
    S.addf0[0] = f->est_k2 != 0; S.addf0[1] = f->est_k4 != 0; S.addf0[2] = f->est_f != 0;
    S.nraddp = S.addf0[0] + S.addf0[1] + S.addf0[2];
    const int N = S.N, pxs = S.pxs, no = S.no;
    char buf[512];
    float ihw2 = 0.f;
    for (int i = 0; i < N; ++i) {
        const int w = io->img_w ? io->img_w[i] : 1, hh = io->img_h ? io->img_h[i] : 1;
        ihw2 += (float)((hh + w) * (hh + w)) * 0.25f;
    }
    S.ihw2 = ihw2 / (float)N;
    S.defdiff = 5000. / (double)sqrtf(S.ihw2);
    std::snprintf(buf, sizeof buf, "#threads: %d %d\n", P, (int)std::lround((double)((float)N / 70.f)));
    S.logs(buf);
    // the partition (rpbacore.cpp:146-150)
    const auto t0 = std::chrono::steady_clock::now();
    S.part.assign((size_t)N, 0);
    int rc = partition_images(N, no, io->obs_cam, io->obs_pt, pxs, P, S.part.data());
    if (rc) return rc;
    std::snprintf(buf, sizeof buf, "\n\nPartitioning runtime: %g s\n\n",
                  std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    S.logs(buf);
    S.pcnt.assign((size_t)P + 1, 0);
    for (int i = 0; i < N; ++i) S.pcnt[(size_t)S.part[i] + 1]++;
    for (int p = 0; p < P; ++p) {
        if (S.pcnt[(size_t)p + 1] == 0) return RPBA_E_INVALID;
        S.pcnt[(size_t)p + 1] += S.pcnt[p];
    }
    S.pimg.assign((size_t)N, 0);
    {
        std::vector<int> pos(S.pcnt.begin(), S.pcnt.end() - 1);
        for (int i = 0; i < N; ++i) S.pimg[(size_t)pos[S.part[i]]++] = i;
    }
    for (int p = 0; p < P; ++p) {
        std::snprintf(buf, sizeof buf, "Thread: %d #images %d indices %d %d\n", p, S.pcnt[(size_t)p + 1] - S.pcnt[p],
                      S.pcnt[p], S.pcnt[(size_t)p + 1]);
        S.logs(buf);
    }
    // tracks in file order (InputBlock, rpba.cpp:219-238)
    S.base.assign((size_t)pxs + 1, 0);
    S.len.assign((size_t)pxs, 0);
    for (int q = 0; q < no; ++q) {
        const int c = io->obs_cam[q], p = io->obs_pt[q];
        if (c < 0) continue;
        if (c >= N || p < 0 || p >= pxs) return RPBA_E_INVALID;
        S.len[p]++;
    }
    for (int i = 0; i < pxs; ++i) S.base[(size_t)i + 1] = S.base[i] + S.len[i];
    S.trk.assign((size_t)std::max<long long>(S.base[pxs], 1), 0);
    std::fill(S.len.begin(), S.len.end(), 0);
    for (int q = 0; q < no; ++q)
        if (io->obs_cam[q] >= 0) {
            const int p = io->obs_pt[q];
            S.trk[(size_t)(S.base[p] + S.len[p]++)] = q;
        }
    // a camera seen twice in one track is refused, as in the global solve;
    // slots: the parts of each point's initial track, ascending (counted over
    // ranges of points, then filled)
    S.soff.assign((size_t)pxs + 1, 0);
    {
        const int nr = S.nthr;
        std::vector<char> dup((size_t)nr, 0);
        auto parts_of = [&](long long i, std::vector<int> &tmp) {
            tmp.clear();
            for (int q = 0; q < S.len[i]; ++q) tmp.push_back(S.part[io->obs_cam[S.trk[(size_t)(S.base[i] + q)]]]);
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        };
        parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
            std::vector<int> seen((size_t)N, -1), tmp;
            for (long long i = i0; i < i1; ++i) {
                for (int q = 0; q < S.len[i]; ++q) {
                    const int c = io->obs_cam[S.trk[(size_t)(S.base[i] + q)]];
                    if (seen[c] == (int)i) dup[r] = 1;
                    seen[c] = (int)i;
                }
                parts_of(i, tmp);
                S.soff[(size_t)i + 1] = (long long)tmp.size();
            }
        });
        for (int r = 0; r < nr; ++r)
            if (dup[r]) return RPBA_E_INVALID;
        for (int i = 0; i < pxs; ++i) S.soff[(size_t)i + 1] += S.soff[i];
        S.spart.resize((size_t)S.soff[pxs]);
        parallel_ranges(pxs, nr, [&](long long i0, long long i1, int) {
            std::vector<int> tmp;
            for (long long i = i0; i < i1; ++i) {
                parts_of(i, tmp);
                std::copy(tmp.begin(), tmp.end(), S.spart.begin() + S.soff[i]);
            }
        });
    }
    const size_t NS = (size_t)S.soff[pxs];
    S.pit.assign(std::max<size_t>(NS, 1), 0);
    S.bkept.assign(std::max<size_t>(NS, 1), 0);
    S.w3h.assign(6 * NS, 0.f);
    S.gcpinfo.assign((size_t)pxs, -1);
    S.sig0wv.assign((size_t)N, 0.f);
    S.XXd.resize(3 * (size_t)pxs);
    for (size_t j = 0; j < S.XXd.size(); ++j) S.XXd[j] = (double)io->pts[j];
    S.PMd.resize(12 * (size_t)N);
    for (size_t j = 0; j < S.PMd.size(); ++j) S.PMd[j] = (double)io->cam_P[j];
    S.cf.assign(io->cam_f, io->cam_f + N);
    S.cd.assign(io->cam_d, io->cam_d + 2 * (size_t)N);
    S.gcpc.assign(3 * (size_t)std::max(pxs, 1), 0.);
    S.w3list.assign((size_t)P, std::vector<float>());
    // device
    if (hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking) != hipSuccess) return RPBA_E_HIP;
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{S.st};
    const size_t ntr0 = (size_t)std::max<long long>(S.base[pxs], 1);
    rc = 0;
    rc |= S.d_base.put(S.base.data(), S.base.size(), S.st);
    rc |= S.d_soff.put(S.soff.data(), S.soff.size(), S.st);
    rc |= S.d_obs_cam.put(io->obs_cam, std::max(no, 1), S.st);
    rc |= S.d_oxy.put(io->obs_xy, 2 * (size_t)std::max(no, 1), S.st);
    if (io->obs_w) rc |= S.d_ow.put(io->obs_w, 3 * (size_t)std::max(no, 1), S.st);
    rc |= S.d_part.put(S.part.data(), S.part.size(), S.st);
    rc |= S.d_spart.put(S.spart.empty() ? S.part.data() : S.spart.data(), std::max<size_t>(NS, 1), S.st);
    rc |= S.d_w3.alloc(6 * std::max<size_t>(NS, 1));
    rc |= S.d_nv.alloc(9 * std::max<size_t>(NS, 1));
    rc |= S.d_wk.alloc(3 * ntr0);
    rc |= S.d_dl.alloc(ntr0);
    rc |= S.d_keys.alloc(ntr0);
    rc |= S.d_nout.alloc((size_t)std::max(pxs, 1));
    rc |= S.d_sg.alloc((size_t)std::max(pxs, 1));
    rc |= S.d_rd.alloc((size_t)std::max(pxs, 1));
    if (rc) return RPBA_E_NOMEM;
    HIPCHK(hipMemsetAsync(S.d_w3.p, 0, sizeof(float) * 6 * std::max<size_t>(NS, 1), S.st));

    // the saved best state (rpbacore.cpp:405-411)
    std::vector<double> XXdold = S.XXd, PMdold = S.PMd;
    std::vector<float> cfold = S.cf, cdold = S.cd;
    std::vector<int> trkold = S.trk, lenold = S.len;
    const int nrparimages = (N - 2) * 6 + 5 + N * S.nraddp;
    double sig0opt = 1.e20, sig0g = 0.;
    int iflag = S.robust ? 2 : 1, iter = 1, nit = 0;
    float sig0frac = 1.01f;
    std::vector<double> sgv;
    std::vector<int> rdv;
    std::vector<float> keys;
    if (stats) {
        stats->passes_done = 0;
        stats->trace_len = 0;
    }
    const int nrppt = pxs / P;
    while (iflag > 0 && iter < 51) {
        int robustit = 0, pcg_it = 0;
        S.cur_iter = iter;
        S.clk.mark(nullptr, "start", iter);
        if (iter > 1) {
            int addf[3] = {S.addf0[0], S.addf0[1], S.addf0[2]};
            if (iter < 3) addf[0] = addf[1] = addf[2] = 0;
            rc = S.block_step(iter, addf, pcg_it);
            if (rc) return rc;
            S.pcg_total += pcg_it;
        }
        if (iflag == 1) {
            robustit = S.robust;
            if (S.robust) sig0frac = 1.02f;
        }
        if (iter > 1) {
            // pointXXv from the parts' tracks, part by part (rpbacore.cpp:276-289);
            // the slots of the entries are those block_step indexed
            parallel_ranges(pxs, S.nthr, [&](long long i0, long long i1, int) {
                std::vector<int> tmp;
                for (long long i = i0; i < i1; ++i) {
                    if (!S.len[i]) continue;
                    tmp.clear();
                    const long long b = S.base[i];
                    for (long long sl = S.soff[i]; sl < S.soff[(size_t)i + 1]; ++sl) {
                        if (!S.bkept[(size_t)sl]) continue;
                        for (int q = 0; q < S.len[i]; ++q)
                            if (S.eslot[(size_t)(b + q)] == sl) tmp.push_back(S.trk[(size_t)(b + q)]);
                    }
                    std::copy(tmp.begin(), tmp.end(), S.trk.begin() + b);
                    S.len[i] = tmp.size() < 2 ? 0 : (int)tmp.size();
                }
            });
        }
        S.clk.mark(nullptr, "tracks regrouped", iter);
        // pitflags of the points with entries (rpbacore.cpp:301-309)
        S.entry_index();
        parallel_ranges(pxs, S.nthr, [&](long long i0, long long i1, int) {
            for (long long i = i0; i < i1; ++i) {
                if (!S.len[i]) continue;
                for (long long sl = S.soff[i]; sl < S.soff[(size_t)i + 1]; ++sl) S.pit[(size_t)sl] = 0;
                for (int q = 0; q < S.len[i]; ++q) S.pit[(size_t)S.eslot[(size_t)(S.base[i] + q)]] = 1;
            }
        });
        S.clk.mark(nullptr, "pitflags", iter);
        if (robustit) {
            // the robust statistics pass and the per-image scales (rpbacore.cpp:311-375)
            std::vector<int> len0 = S.len;
            rc = S.point_pass(1, 0, 0, sgv, rdv, &keys);
            if (rc) return rc;
            // the keys grouped by image (a counting sort over ranges of
            // points), then each image's upper median, which does not depend
            // on the order inside the group
            const int nr = S.nthr;
            std::vector<long long> kc((size_t)nr * N + 1, 0);
            parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
                long long *c = kc.data() + (size_t)r * N;
                for (long long i = i0; i < i1; ++i)
                    for (int q = 0; q < len0[i]; ++q) c[io->obs_cam[S.trk[(size_t)(S.base[i] + q)]]]++;
            });
            std::vector<long long> koff((size_t)N + 1, 0);
            {
                long long at = 0;
                for (int k = 0; k < N; ++k) {
                    koff[k] = at;
                    for (int r = 0; r < nr; ++r) {
                        const long long n = kc[(size_t)r * N + k];
                        kc[(size_t)r * N + k] = at;
                        at += n;
                    }
                }
                koff[N] = at;
            }
            std::vector<float> kv((size_t)std::max<long long>(koff[N], 1));
            parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
                long long *c = kc.data() + (size_t)r * N;
                for (long long i = i0; i < i1; ++i)
                    for (int q = 0; q < len0[i]; ++q) {
                        const long long e = S.base[i] + q;
                        kv[(size_t)c[io->obs_cam[S.trk[(size_t)e]]]++] = keys[(size_t)e];
                    }
            });
            parallel_ranges(N, nr, [&](long long k0, long long k1, int) {
                for (long long k = k0; k < k1; ++k) {
                    float *v0 = kv.data() + koff[k], *v1 = kv.data() + koff[(size_t)k + 1];
                    if (v1 == v0) continue;
                    const size_t n = (size_t)(v1 - v0);
                    std::nth_element(v0, v0 + n / 2, v1);
                    S.sig0wv[k] = 1.f / (v0[n / 2] * 2.1981f);
                }
            });
            for (int k = 0; k < N; ++k)
                if (koff[(size_t)k + 1] == koff[k]) {
                    std::snprintf(buf, sizeof buf, "Image %d without observations!!!!!!!!!!!!!!!!!!!!!\n\n", k);
                    S.logs(buf);
                    S.sig0wv[k] = 0.f;
                }
        }
        S.clk.mark(nullptr, "robust statistics pass", iter);
        // the main point pass (rpbacore.cpp:378-392)
        rc = S.point_pass(robustit ? 2 : 0, iter == 1 ? 1 : 0, 1, sgv, rdv, nullptr);
        if (rc) return rc;
        S.clk.mark(nullptr, "main point pass", iter);
        sig0g = 0.;
        long redundancy = 0;
        for (int t = 0; t < P; ++t) {
            const int b0 = nrppt * t, b1 = (t == P - 1) ? pxs : nrppt * (t + 1);
            double s = 0.;
            long r = 0;
            for (int i = b0; i < b1; ++i) {
                s += sgv[i];
                r += rdv[i];
            }
            sig0g += s;
            redundancy += r;
        }
        long long nrobs = 0;
        for (int i = 0; i < pxs; ++i) nrobs += S.len[i];
        nrobs *= 2;
        sig0g /= (double)(redundancy - nrparimages);
        std::snprintf(buf, sizeof buf, "IT: %d %d  %g %g  %g %ld %lld %d %d  %ld\n", iter, iflag,
                      std::sqrt(sig0g * (double)S.ihw2), sig0opt / sig0g, sig0g, redundancy, nrobs, nrparimages, S.nraddp,
                      redundancy - nrparimages);
        S.logs(buf);
        const bool saved = sig0g <= sig0opt;
        if (saved) {
            parallel_assign(XXdold, S.XXd, S.nthr);
            PMdold = S.PMd;
            parallel_assign(trkold, S.trk, S.nthr);
            parallel_assign(lenold, S.len, S.nthr);
            cfold = S.cf;
            cdold = S.cd;
        }
        if (stats && stats->trace_sig0 && stats->trace_len < stats->trace_cap) {
            stats->trace_sig0[stats->trace_len] = sig0g;
            if (stats->trace_accept) stats->trace_accept[stats->trace_len] = saved ? 1 : 0;
            if (stats->trace_pcg) stats->trace_pcg[stats->trace_len] = pcg_it;
            if (stats->trace_nrobs) stats->trace_nrobs[stats->trace_len] = (int)nrobs;
            stats->trace_len++;
        }
        // GCPs: points with entries in two parts or more, their centres and
        // each part's weight list (rpbacore.cpp:419-450)
        {
            // GCP numbers in point order and each part's list in point
            // order: counts over ranges of points, then the ranges fill
            // their own places
            const int nr = S.nthr;
            auto is_gcp = [&](long long i) {
                if (S.len[i] < 2) return false;
                int gnum = 0;
                for (long long s = S.soff[i]; s < S.soff[(size_t)i + 1] && gnum < 2; ++s) gnum += S.pit[(size_t)s];
                return gnum > 1;
            };
            std::vector<long long> gc((size_t)nr * (P + 1), 0);  // per range: GCP points, then entries per part
            parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
                long long *c = gc.data() + (size_t)r * (P + 1);
                for (long long i = i0; i < i1; ++i) {
                    if (!is_gcp(i)) continue;
                    c[0]++;
                    for (long long s = S.soff[i]; s < S.soff[(size_t)i + 1]; ++s)
                        if (S.pit[(size_t)s]) c[1 + S.spart[(size_t)s]]++;
                }
            });
            for (int j = 0; j <= P; ++j) {
                long long at = 0;
                for (int r = 0; r < nr; ++r) {
                    const long long n = gc[(size_t)r * (P + 1) + j];
                    gc[(size_t)r * (P + 1) + j] = at;
                    at += n;
                }
                if (j > 0) S.w3list[(size_t)j - 1].assign(6 * (size_t)at, 0.f);
            }
            parallel_ranges(pxs, nr, [&](long long i0, long long i1, int r) {
                long long *c = gc.data() + (size_t)r * (P + 1);
                for (long long i = i0; i < i1; ++i) {
                    if (!is_gcp(i)) {
                        S.gcpinfo[i] = -1;
                        continue;
                    }
                    const long long g = c[0]++;
                    S.gcpinfo[i] = (int)g;
                    for (int j = 0; j < 3; ++j) S.gcpc[3 * (size_t)g + j] = S.XXd[3 * (size_t)i + j];
                    for (long long s = S.soff[i]; s < S.soff[(size_t)i + 1]; ++s) {
                        if (!S.pit[(size_t)s]) continue;
                        const int p = S.spart[(size_t)s];
                        float *L = S.w3list[(size_t)p].data() + 6 * (size_t)(c[1 + p]++);
                        for (int v = 0; v < 6; ++v) L[v] = S.w3h[6 * (size_t)s + v];
                    }
                }
            });
        }
        S.clk.mark(nullptr, "sums, best state, GCP lists", iter);
        // iflag (rpbacore.cpp:479-493)
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
    // the outputs of the saved state (rpbacore.cpp:497-507)
    if (io->obs_kept) std::memset(io->obs_kept, 0, (size_t)no);
    if (io->pt_kept) std::memset(io->pt_kept, 0, (size_t)pxs);
    long long nobs_k = 0;
    int npts_k = 0;
    double ss = 0.;
    for (int i = 0; i < pxs; ++i) {
        for (int j = 0; j < 3; ++j) io->pts[3 * (size_t)i + j] = (float)XXdold[3 * (size_t)i + j];
        if (lenold[i] > 0) {
            ++npts_k;
            if (io->pt_kept) io->pt_kept[i] = 1;
        }
        for (int q = 0; q < lenold[i]; ++q) {
            const int fo = trkold[(size_t)(S.base[i] + q)], k = io->obs_cam[fo];
            if (io->obs_kept) io->obs_kept[fo] = 1;
            ++nobs_k;
            // RMS of the output state, double
            const double *Pk = PMdold.data() + 12 * (size_t)k;
            const double *x = XXdold.data() + 3 * (size_t)i;
            const double e0 = Pk[0] * x[0] + Pk[1] * x[1] + Pk[2] * x[2] + Pk[3];
            const double e1 = Pk[4] * x[0] + Pk[5] * x[1] + Pk[6] * x[2] + Pk[7];
            const double e2 = Pk[8] * x[0] + Pk[9] * x[1] + Pk[10] * x[2] + Pk[11];
            const double u = e0 / e2, v = e1 / e2, rho = u * u + v * v;
            const double sc = 1. + (double)cdold[2 * (size_t)k] * rho + (double)cdold[2 * (size_t)k + 1] * rho * rho;
            const double r0 = (double)io->obs_xy[2 * (size_t)fo] - (double)cfold[k] * sc * u;
            const double r1 = (double)io->obs_xy[2 * (size_t)fo + 1] - (double)cfold[k] * sc * v;
            ss += r0 * r0 + r1 * r1;
        }
    }
    for (size_t j = 0; j < 12 * (size_t)N; ++j) io->cam_P[j] = (float)PMdold[j];
    std::memcpy(io->cam_f, cfold.data(), sizeof(float) * (size_t)N);
    std::memcpy(io->cam_d, cdold.data(), sizeof(float) * 2 * (size_t)N);
    if (stats) {
        stats->itt = nit;
        stats->pcg_iters = S.pcg_total;
        stats->sig0_final = sig0opt;
        stats->passes_done = 1;
        stats->n_obs_kept = nobs_k;
        stats->n_pts_kept = npts_k;
        stats->rms_px = nobs_k ? std::sqrt(ss / (double)(2 * nobs_k)) : 0.;
    }
    return RPBA_OK;
}

}  // namespace rpba

extern "C" int rpba_partition_images(int32_t n_cams, int32_t n_obs, const int32_t *obs_cam, const int32_t *obs_pt,
                                     int32_t n_pts, int32_t n_parts, int32_t *part)
{
    return rpba::partition_images(n_cams, n_obs, obs_cam, obs_pt, n_pts, n_parts, part);
}
