// Learned correction of the block-Jacobi preconditioner: graph features from
// S, the graph network's forward and backward passes, Adam, and the C ABI of
// training and inference (include/rpba.h, rpba_lp_*).  Every reduction runs
// in a fixed order (per-node loops in edge order, per-chunk partials summed in
// chunk order), so training is bitwise reproducible.
#include "lp.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../include/rpba.h"

namespace rpba {

namespace {

constexpr int H = kLpH, F = kLpF;
constexpr int kChunk = 256;  // nodes per weight-gradient partial
// activation slots (N x H BF16 each): h0, pre_in, then per layer m, pre, h_next
constexpr int kSlots = 2 + 3 * kLpL;
__host__ __device__ constexpr int slot_h(int i) { return i == 0 ? 0 : 4 + 3 * (i - 1); }
__host__ __device__ constexpr int slot_m(int i) { return 2 + 3 * i; }
__host__ __device__ constexpr int slot_pre(int i) { return 3 + 3 * i; }
constexpr int kSlotPreIn = 1;
// parameter offsets
constexpr int oWin = 0, oBin = H * F;
__host__ __device__ constexpr int oWs(int i) { return H * F + H + i * (2 * H * H + H); }
__host__ __device__ constexpr int oWn(int i) { return oWs(i) + H * H; }
__host__ __device__ constexpr int oB(int i) { return oWs(i) + 2 * H * H; }
constexpr int oV = H * F + H + kLpL * (2 * H * H + H), oC = oV + H;
static_assert(oC + 1 == kLpParams, "parameter layout");

__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// graph of S
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(64) void k_lp_norms(int nc, const double *Sd, float *nrm)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    double s = 0.;
    for (int j = 0; j < 81; ++j) s += Sd[81 * (size_t)k + j] * Sd[81 * (size_t)k + j];
    nrm[k] = (float)sqrt(s);
}

__global__ __launch_bounds__(256) void k_lp_edges(int nblk, const double *Su, SuMap sm, const int *bk, const int *bl,
                                                  const float *nrm, float *eb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    double s = 0.;
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
            const double v = Su[su_at(sm, b, i, j)];
            s += v * v;
        }
    const double den = sqrt((double)nrm[bk[b]] * (double)nrm[bl[b]]);
    eb[b] = den > 0. ? (float)(sqrt(s) / den) : 0.f;
}

// node k: its neighbours in row order (blocks (k, l), l > k) then column
// order (blocks (l, k), l < k), and its 48 features
__global__ __launch_bounds__(64) void k_lp_csr(int nc, const double *Sd, const int *bk, const int *bl,
                                               const int *row_off, const int *col_off, const int *col_blk,
                                               const float *eb, float *feat, int *adj_off, int *adj_nbr,
                                               float *adj_w)
{
    const int k = blockIdx.x, t = threadIdx.x;
    const int r0 = row_off[k], r1 = row_off[k + 1], c0 = col_off[k], c1 = col_off[k + 1];
    const int base = r0 + c0, nr = r1 - r0, deg = nr + (c1 - c0);
    if (t == 0) {
        adj_off[k] = base;
        if (k == nc - 1) adj_off[nc] = base + deg;
    }
    for (int e = t; e < deg; e += 64) {
        int b, nb;
        if (e < nr) {
            b = r0 + e;
            nb = bl[b];
        } else {
            b = col_blk[c0 + e - nr];
            nb = bk[b];
        }
        adj_nbr[base + e] = nb;
        adj_w[base + e] = eb[b];
    }
    __syncthreads();
    if (t == 0) {
        float s = 0.f, mx = 0.f;
        for (int e = 0; e < deg; ++e) {
            const float w = adj_w[base + e];
            s += w;
            mx = fmaxf(mx, w);
        }
        const double *D = Sd + 81 * (size_t)k;
        float ld[9], mean = 0.f;
        for (int i = 0; i < 9; ++i) {
            ld[i] = (float)log(fmax(D[10 * i], 1e-300));
            mean += ld[i];
        }
        mean /= 9.f;
        float *f = feat + (size_t)k * F;
        for (int i = 0; i < 9; ++i) f[i] = ld[i] - mean;
        int n = 9;
        for (int i = 0; i < 9; ++i)
            for (int j = i + 1; j < 9; ++j) {
                const double den = sqrt(fmax(D[10 * i], 1e-300) * fmax(D[10 * j], 1e-300));
                f[n++] = (float)(D[9 * i + j] / den);
            }
        f[45] = logf(1.f + (float)deg) / 8.f;
        f[46] = logf(1.f + s);
        f[47] = mx;
    }
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

__device__ inline uint16_t *slot(uint16_t *act, int nn, int s) { return act + (size_t)s * nn * H; }

// per node: sum of its edge weights, in edge order
__global__ __launch_bounds__(256) void k_lp_nodesum(LpGraphs G, float *nsum)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= G.nn) return;
    float s = 0.f;
    for (int e = G.adj_off[k]; e < G.adj_off[k + 1]; ++e) s += G.adj_w[e];
    nsum[k] = s;
}

__global__ __launch_bounds__(64) void k_lp_input(LpGraphs G, const uint16_t *pb, uint16_t *act)
{
    const int k = blockIdx.x, o = threadIdx.x;
    __shared__ float x[F];
    if (o < F) x[o] = bf2f(f2bf(G.feat[(size_t)k * F + o]));
    __syncthreads();
    float s = bf2f(pb[oBin + o]);
    const uint16_t *w = pb + oWin + o * F;
    for (int j = 0; j < F; ++j) s += bf2f(w[j]) * x[j];
    const uint16_t pre = f2bf(s);
    slot(act, G.nn, kSlotPreIn)[(size_t)k * H + o] = pre;
    slot(act, G.nn, 0)[(size_t)k * H + o] = f2bf(fmaxf(bf2f(pre), 0.f));
}

// m_k = (sum_e w_e h_nbr(e)) / sum_e w_e, in edge order
__global__ __launch_bounds__(64) void k_lp_aggregate(LpGraphs G, const float *nsum, uint16_t *act, int layer)
{
    const int k = blockIdx.x, o = threadIdx.x;
    const uint16_t *h = slot(act, G.nn, slot_h(layer));
    float s = 0.f;
    for (int e = G.adj_off[k]; e < G.adj_off[k + 1]; ++e) s += G.adj_w[e] * bf2f(h[(size_t)G.adj_nbr[e] * H + o]);
    const float ns = nsum[k];
    slot(act, G.nn, slot_m(layer))[(size_t)k * H + o] = f2bf(ns > 0.f ? s / ns : 0.f);
}

__global__ __launch_bounds__(64) void k_lp_layer(LpGraphs G, const uint16_t *pb, uint16_t *act, int layer)
{
    const int k = blockIdx.x, o = threadIdx.x;
    __shared__ float hv[H], mv[H];
    hv[o] = bf2f(slot(act, G.nn, slot_h(layer))[(size_t)k * H + o]);
    mv[o] = bf2f(slot(act, G.nn, slot_m(layer))[(size_t)k * H + o]);
    __syncthreads();
    float s = bf2f(pb[oB(layer) + o]);
    const uint16_t *ws = pb + oWs(layer) + o * H, *wn = pb + oWn(layer) + o * H;
    for (int j = 0; j < H; ++j) s += bf2f(ws[j]) * hv[j];
    for (int j = 0; j < H; ++j) s += bf2f(wn[j]) * mv[j];
    const uint16_t pre = f2bf(s);
    slot(act, G.nn, slot_pre(layer))[(size_t)k * H + o] = pre;
    slot(act, G.nn, slot_h(layer + 1))[(size_t)k * H + o] = f2bf(hv[o] + fmaxf(bf2f(pre), 0.f));
}

// per graph: g = mean of h, w = wmax sigmoid(v . g + c); loss share
__global__ __launch_bounds__(64) void k_lp_readout(LpGraphs G, const uint16_t *pb, const uint16_t *act,
                                                   float *gvec, float *omega, float *loss)
{
    const int g = blockIdx.x, o = threadIdx.x;
    __shared__ float gv[H];
    const uint16_t *h = act + (size_t)slot_h(kLpL) * G.nn * H;
    const int n0 = G.goff[g], n1 = G.goff[g + 1];
    float s = 0.f;
    for (int k = n0; k < n1; ++k) s += bf2f(h[(size_t)k * H + o]);
    s = (n1 > n0) ? s / (float)(n1 - n0) : 0.f;
    gv[o] = s;
    gvec[(size_t)g * H + o] = s;
    __syncthreads();
    if (o == 0) {
        float z = bf2f(pb[oC]);
        for (int j = 0; j < H; ++j) z += bf2f(pb[oV + j]) * gv[j];
        const float w = kLpWmax / (1.f + expf(-z));
        omega[g] = w;
        if (loss) loss[g] = G.label ? (w - G.label[g]) * (w - G.label[g]) : 0.f;
    }
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------

// dL/dw = 2 (w - y) / ng; dz = dL/dw w (1 - w / wmax); per graph the shares
// of dv (gpart row g) and dc, and dh = dz v / n on the graph's nodes
__global__ __launch_bounds__(64) void k_lp_readout_bwd(LpGraphs G, const uint16_t *pb, const float *gvec,
                                                       const float *omega, float *dh, float *gpart)
{
    const int g = blockIdx.x, o = threadIdx.x;
    const float w = omega[g];
    const float dw = 2.f * (w - G.label[g]) / (float)G.ng;
    const float dz = dw * w * (1.f - w / kLpWmax);
    gpart[(size_t)g * (H + 1) + o] = dz * gvec[(size_t)g * H + o];
    if (o == 0) gpart[(size_t)g * (H + 1) + H] = dz;
    const int n0 = G.goff[g], n1 = G.goff[g + 1];
    const float dv = (n1 > n0) ? dz * bf2f(pb[oV + o]) / (float)(n1 - n0) : 0.f;
    for (int k = n0; k < n1; ++k) dh[(size_t)k * H + o] = dv;
}

// layer backward at node k: dpre = dh (pre > 0); q = W_n^T dpre;
// dh <- dh + W_s^T dpre (the residual path keeps dh)
__global__ __launch_bounds__(64) void k_lp_layer_bwd(LpGraphs G, const uint16_t *pb, const uint16_t *act, float *dh,
                                                     float *dpre, float *q, int layer)
{
    const int k = blockIdx.x, o = threadIdx.x;
    __shared__ float dp[H];
    const float pre = bf2f(slot(const_cast<uint16_t *>(act), G.nn, slot_pre(layer))[(size_t)k * H + o]);
    const float d = dh[(size_t)k * H + o];
    const float p = pre > 0.f ? d : 0.f;
    dp[o] = p;
    dpre[(size_t)k * H + o] = p;
    __syncthreads();
    float sq = 0.f, ss = 0.f;
    for (int j = 0; j < H; ++j) {
        sq += bf2f(pb[oWn(layer) + j * H + o]) * dp[j];
        ss += bf2f(pb[oWs(layer) + j * H + o]) * dp[j];
    }
    q[(size_t)k * H + o] = sq;
    dh[(size_t)k * H + o] = d + ss;
}

// dh_l += sum_e (w_e / nsum_nbr(e)) q_nbr(e), in edge order (the adjacency is
// symmetric with equal weights both ways)
__global__ __launch_bounds__(64) void k_lp_aggregate_T(LpGraphs G, const float *nsum, const float *q, float *dh)
{
    const int l = blockIdx.x, o = threadIdx.x;
    float s = 0.f;
    for (int e = G.adj_off[l]; e < G.adj_off[l + 1]; ++e) {
        const int k = G.adj_nbr[e];
        const float ns = nsum[k];
        if (ns > 0.f) s += (G.adj_w[e] / ns) * q[(size_t)k * H + o];
    }
    dh[(size_t)l * H + o] += s;
}

// input layer: dpre_in = dh (pre_in > 0)
__global__ __launch_bounds__(64) void k_lp_input_bwd(LpGraphs G, const uint16_t *act, const float *dh, float *dpre)
{
    const int k = blockIdx.x, o = threadIdx.x;
    const float pre = bf2f(act[(size_t)kSlotPreIn * G.nn * H + (size_t)k * H + o]);
    dpre[(size_t)k * H + o] = pre > 0.f ? dh[(size_t)k * H + o] : 0.f;
}

// weight-gradient partials over a chunk of nodes: part[c][o][j] =
// sum_k A[k][o] B[k][j] (j < nb), part[c][o][nb] = sum_k A[k][o]; B is BF16
// activations (slot) or, with feat, the BF16-rounded features
__global__ __launch_bounds__(256) void k_lp_wgrad(int nn, const float *A, const uint16_t *B, const float *feat, int nb,
                                                  float *part)
{
    const int c = blockIdx.x;
    const int k0 = c * kChunk, k1 = min(nn, k0 + kChunk);
    const int nout = H * (nb + 1);
    for (int e = threadIdx.x; e < nout; e += blockDim.x) {
        const int o = e / (nb + 1), j = e - o * (nb + 1);
        float s = 0.f;
        for (int k = k0; k < k1; ++k) {
            const float bv = (j == nb) ? 1.f : (feat ? bf2f(f2bf(feat[(size_t)k * nb + j])) : bf2f(B[(size_t)k * nb + j]));
            s += A[(size_t)k * H + o] * bv;
        }
        part[(size_t)c * nout + e] = s;
    }
}

// gradient <- sum over chunks in order; W part to gW [H][nb], bias to gb [H]
__global__ __launch_bounds__(256) void k_lp_wreduce(int nchunk, int nb, const float *part, float *gW, float *gb)
{
    const int nout = H * (nb + 1);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nout) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * nout + e];
    const int o = e / (nb + 1), j = e - o * (nb + 1);
    if (j == nb) {
        if (gb) gb[o] = s;
    } else {
        gW[o * nb + j] = s;
    }
}

// readout gradient: dv and dc summed over graphs in order
__global__ __launch_bounds__(128) void k_lp_vreduce(int ng, const float *gpart, float *gv)
{
    const int o = threadIdx.x;
    if (o > H) return;
    float s = 0.f;
    for (int g = 0; g < ng; ++g) s += gpart[(size_t)g * (H + 1) + o];
    gv[o] = s;  // o == H: dc
}

__global__ __launch_bounds__(256) void k_lp_adam(int n, float *p, const float *g, float *m, float *v, uint16_t *pb,
                                                 float lr, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    p[i] = pi;
    pb[i] = f2bf(pi);
}

__global__ __launch_bounds__(256) void k_lp_tobf(int n, const float *p, uint16_t *pb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = f2bf(p[i]);
}

template <class Tp>
int dalloc(Tp **p, size_t n)
{
    *p = nullptr;
    return hipMalloc((void **)p, sizeof(Tp) * std::max<size_t>(n, 1)) == hipSuccess ? 0 : RPBA_E_NOMEM;
}
template <class Tp>
void dfree(Tp *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

void lp_graph_from_S(int nc, int nblk, const double *Sd, const double *Su, const SuMap &sm, const int *bk, const int *bl,
                     const int *row_off, const int *col_off, const int *col_blk, float *nrm, float *eb, float *feat,
                     int *adj_off, int *adj_nbr, float *adj_w, hipStream_t st)
{
    if (nc <= 0) return;
    k_lp_norms<<<(nc + 63) / 64, 64, 0, st>>>(nc, Sd, nrm);
    if (nblk > 0) k_lp_edges<<<(nblk + 255) / 256, 256, 0, st>>>(nblk, Su, sm, bk, bl, nrm, eb);
    k_lp_csr<<<nc, 64, 0, st>>>(nc, Sd, bk, bl, row_off, col_off, col_blk, eb, feat, adj_off, adj_nbr, adj_w);
}

LpModel::~LpModel()
{
    dfree(p); dfree(pb); dfree(g); dfree(m); dfree(v); dfree(gpart); dfree(act); dfree(dh); dfree(dpre); dfree(q);
    dfree(nsum); dfree(gvec); dfree(omega); dfree(loss);
}

int LpModel::init(const float *params, uint64_t seed)
{
    int rc = 0;
    rc |= dalloc(&p, kLpParams);
    rc |= dalloc(&pb, kLpParams);
    rc |= dalloc(&g, kLpParams);
    rc |= dalloc(&m, kLpParams);
    rc |= dalloc(&v, kLpParams);
    if (rc) return RPBA_E_NOMEM;
    std::vector<float> h(kLpParams, 0.f);
    if (params) {
        std::memcpy(h.data(), params, sizeof(float) * kLpParams);
    } else {
        // He-style normal initialisation from the seed; biases 0; the
        // readout starts at w = wmax sigmoid(0.5)
        std::mt19937_64 gen(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        for (int i = 0; i < H * F; ++i) h[oWin + i] = nd(gen) * std::sqrt(2.f / F);
        for (int l = 0; l < kLpL; ++l) {
            for (int i = 0; i < H * H; ++i) h[oWs(l) + i] = nd(gen) * std::sqrt(1.f / H) * 0.5f;
            for (int i = 0; i < H * H; ++i) h[oWn(l) + i] = nd(gen) * std::sqrt(1.f / H) * 0.5f;
        }
        for (int i = 0; i < H; ++i) h[oV + i] = nd(gen) * std::sqrt(1.f / H) * 0.1f;
        h[oC] = 0.5f;
    }
    if (hipMemcpy(p, h.data(), sizeof(float) * kLpParams, hipMemcpyHostToDevice) != hipSuccess) return RPBA_E_HIP;
    if (hipMemset(m, 0, sizeof(float) * kLpParams) != hipSuccess) return RPBA_E_HIP;
    if (hipMemset(v, 0, sizeof(float) * kLpParams) != hipSuccess) return RPBA_E_HIP;
    k_lp_tobf<<<(kLpParams + 255) / 256, 256>>>(kLpParams, p, pb);
    adam_t = 0;
    return hipDeviceSynchronize() == hipSuccess ? 0 : RPBA_E_HIP;
}

int LpModel::reserve(int nn, int ng)
{
    if ((size_t)nn > cap_nodes) {
        dfree(act); dfree(dh); dfree(dpre); dfree(q); dfree(nsum);
        int rc = 0;
        rc |= dalloc(&act, (size_t)kSlots * nn * H);
        rc |= dalloc(&dh, (size_t)nn * H);
        rc |= dalloc(&dpre, (size_t)nn * H);
        rc |= dalloc(&q, (size_t)nn * H);
        rc |= dalloc(&nsum, (size_t)nn);
        if (rc) return RPBA_E_NOMEM;
        cap_nodes = nn;
    }
    if ((size_t)ng > cap_graphs) {
        dfree(gvec); dfree(omega); dfree(loss);
        int rc = 0;
        rc |= dalloc(&gvec, (size_t)ng * H);
        rc |= dalloc(&omega, (size_t)ng);
        rc |= dalloc(&loss, (size_t)ng);
        if (rc) return RPBA_E_NOMEM;
        cap_graphs = ng;
    }
    const size_t nchunk = (size_t)(nn + kChunk - 1) / kChunk;
    const size_t need = std::max(nchunk * H * (H + 1), (size_t)ng * (H + 1));
    if (need > gpart_cap) {
        dfree(gpart);
        if (dalloc(&gpart, need)) return RPBA_E_NOMEM;
        gpart_cap = need;
    }
    return 0;
}

void LpModel::forward(const LpGraphs &G, float *omega_out, hipStream_t st)
{
    if (G.nn <= 0 || G.ng <= 0) return;
    k_lp_nodesum<<<(G.nn + 255) / 256, 256, 0, st>>>(G, nsum);
    k_lp_input<<<G.nn, 64, 0, st>>>(G, pb, act);
    for (int l = 0; l < kLpL; ++l) {
        k_lp_aggregate<<<G.nn, 64, 0, st>>>(G, nsum, act, l);
        k_lp_layer<<<G.nn, 64, 0, st>>>(G, pb, act, l);
    }
    k_lp_readout<<<G.ng, 64, 0, st>>>(G, pb, act, gvec, omega_out ? omega_out : omega, loss);
}

void LpModel::forward_backward(const LpGraphs &G, hipStream_t st)
{
    forward(G, omega, st);
    if (G.nn <= 0 || G.ng <= 0 || !G.label) return;
    const int nchunk = (G.nn + kChunk - 1) / kChunk;
    k_lp_readout_bwd<<<G.ng, 64, 0, st>>>(G, pb, gvec, omega, dh, gpart);
    k_lp_vreduce<<<1, 128, 0, st>>>(G.ng, gpart, g + oV);
    for (int l = kLpL - 1; l >= 0; --l) {
        k_lp_layer_bwd<<<G.nn, 64, 0, st>>>(G, pb, act, dh, dpre, q, l);
        k_lp_aggregate_T<<<G.nn, 64, 0, st>>>(G, nsum, q, dh);
        k_lp_wgrad<<<nchunk, 256, 0, st>>>(G.nn, dpre, act + (size_t)slot_h(l) * G.nn * H, nullptr, H, gpart);
        k_lp_wreduce<<<(H * (H + 1) + 255) / 256, 256, 0, st>>>(nchunk, H, gpart, g + oWs(l), g + oB(l));
        k_lp_wgrad<<<nchunk, 256, 0, st>>>(G.nn, dpre, act + (size_t)slot_m(l) * G.nn * H, nullptr, H, gpart);
        k_lp_wreduce<<<(H * (H + 1) + 255) / 256, 256, 0, st>>>(nchunk, H, gpart, g + oWn(l), nullptr);
    }
    k_lp_input_bwd<<<G.nn, 64, 0, st>>>(G, act, dh, dpre);
    k_lp_wgrad<<<nchunk, 256, 0, st>>>(G.nn, dpre, nullptr, G.feat, F, gpart);
    k_lp_wreduce<<<(H * (F + 1) + 255) / 256, 256, 0, st>>>(nchunk, F, gpart, g + oWin, g + oBin);
}

void LpModel::adam(float lr, hipStream_t st)
{
    ++adam_t;
    const float c1 = 1.f - std::pow(0.9f, (float)adam_t), c2 = 1.f - std::pow(0.999f, (float)adam_t);
    k_lp_adam<<<(kLpParams + 255) / 256, 256, 0, st>>>(kLpParams, p, g, m, v, pb, lr, c1, c2);
}

}  // namespace rpba

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

using namespace rpba;

struct rpba_lp {
    LpModel model;
    hipStream_t st = nullptr;
    // device copy of the last host graph set
    int *goff = nullptr, *adj_off = nullptr, *adj_nbr = nullptr;
    float *feat = nullptr, *adj_w = nullptr, *label = nullptr;
    size_t cap_g = 0, cap_n = 0, cap_e = 0;
};

static int lp_upload(rpba_lp *m, const rpba_lp_graphs *gs, LpGraphs &G)
{
    if (!gs || gs->n_graphs < 1 || gs->n_nodes < 1 || gs->n_edges < 0 || !gs->goff || !gs->feat || !gs->adj_off ||
        (gs->n_edges > 0 && (!gs->adj_nbr || !gs->adj_w)))
        return RPBA_E_INVALID;
    if (gs->goff[0] != 0 || gs->goff[gs->n_graphs] != gs->n_nodes || gs->adj_off[gs->n_nodes] != gs->n_edges)
        return RPBA_E_INVALID;
    for (int64_t e = 0; e < gs->n_edges; ++e)
        if (gs->adj_nbr[e] < 0 || gs->adj_nbr[e] >= gs->n_nodes) return RPBA_E_INVALID;
    const size_t ng = gs->n_graphs, nn = gs->n_nodes, ne = gs->n_edges;
    if (ng > m->cap_g) {
        dfree(m->goff); dfree(m->label);
        if (dalloc(&m->goff, ng + 1) || dalloc(&m->label, ng)) return RPBA_E_NOMEM;
        m->cap_g = ng;
    }
    if (nn > m->cap_n) {
        dfree(m->adj_off); dfree(m->feat);
        if (dalloc(&m->adj_off, nn + 1) || dalloc(&m->feat, nn * kLpF)) return RPBA_E_NOMEM;
        m->cap_n = nn;
    }
    if (ne > m->cap_e) {
        dfree(m->adj_nbr); dfree(m->adj_w);
        if (dalloc(&m->adj_nbr, ne) || dalloc(&m->adj_w, ne)) return RPBA_E_NOMEM;
        m->cap_e = ne;
    }
    bool ok = true;
    ok &= hipMemcpy(m->goff, gs->goff, sizeof(int) * (ng + 1), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(m->adj_off, gs->adj_off, sizeof(int) * (nn + 1), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(m->feat, gs->feat, sizeof(float) * nn * kLpF, hipMemcpyHostToDevice) == hipSuccess;
    if (ne) {
        ok &= hipMemcpy(m->adj_nbr, gs->adj_nbr, sizeof(int) * ne, hipMemcpyHostToDevice) == hipSuccess;
        ok &= hipMemcpy(m->adj_w, gs->adj_w, sizeof(float) * ne, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (gs->label) ok &= hipMemcpy(m->label, gs->label, sizeof(float) * ng, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) return RPBA_E_HIP;
    G.ng = (int)ng;
    G.nn = (int)nn;
    G.ne = (long long)ne;
    G.goff = m->goff;
    G.feat = m->feat;
    G.adj_off = m->adj_off;
    G.adj_nbr = m->adj_nbr;
    G.adj_w = m->adj_w;
    G.label = gs->label ? m->label : nullptr;
    return m->model.reserve((int)nn, (int)ng);
}

extern "C" int rpba_lp_create(rpba_lp **out, const float *params, uint64_t seed)
{
    if (!out) return RPBA_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RPBA_E_HIP;
    rpba_lp *m = new rpba_lp();
    int rc = m->model.init(params, seed);
    if (!rc && hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess) rc = RPBA_E_HIP;
    if (rc) {
        rpba_lp_destroy(m);
        return rc;
    }
    *out = m;
    return RPBA_OK;
}

extern "C" void rpba_lp_destroy(rpba_lp *m)
{
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    dfree(m->goff); dfree(m->adj_off); dfree(m->adj_nbr); dfree(m->feat); dfree(m->adj_w); dfree(m->label);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

extern "C" int rpba_lp_n_params(void) { return kLpParams; }

extern "C" int rpba_lp_get_params(rpba_lp *m, float *out)
{
    if (!m || !out) return RPBA_E_INVALID;
    return hipMemcpy(out, m->model.p, sizeof(float) * kLpParams, hipMemcpyDeviceToHost) == hipSuccess ? RPBA_OK
                                                                                                   : RPBA_E_HIP;
}

extern "C" int rpba_lp_predict(rpba_lp *m, const rpba_lp_graphs *gs, float *omega)
{
    if (!m || !omega) return RPBA_E_INVALID;
    LpGraphs G;
    int rc = lp_upload(m, gs, G);
    if (rc) return rc;
    G.label = nullptr;
    m->model.forward(G, m->model.omega, m->st);
    if (hipMemcpyAsync(omega, m->model.omega, sizeof(float) * G.ng, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess)
        return RPBA_E_HIP;
    return RPBA_OK;
}

extern "C" int rpba_lp_gradient(rpba_lp *m, const rpba_lp_graphs *gs, float *omega, float *loss, float *grad)
{
    if (!m || !gs || !gs->label) return RPBA_E_INVALID;
    LpGraphs G;
    int rc = lp_upload(m, gs, G);
    if (rc) return rc;
    m->model.forward_backward(G, m->st);
    std::vector<float> l(G.ng);
    bool ok = hipMemcpyAsync(l.data(), m->model.loss, sizeof(float) * G.ng, hipMemcpyDeviceToHost, m->st) == hipSuccess;
    if (omega) ok &= hipMemcpyAsync(omega, m->model.omega, sizeof(float) * G.ng, hipMemcpyDeviceToHost, m->st) == hipSuccess;
    if (grad) ok &= hipMemcpyAsync(grad, m->model.g, sizeof(float) * kLpParams, hipMemcpyDeviceToHost, m->st) == hipSuccess;
    ok &= hipStreamSynchronize(m->st) == hipSuccess;
    if (!ok) return RPBA_E_HIP;
    if (loss) {
        double s = 0.;
        for (float x : l) s += x;
        *loss = (float)(s / G.ng);
    }
    return RPBA_OK;
}

extern "C" int rpba_lp_train(rpba_lp *m, const rpba_lp_graphs *gs, int32_t steps, float lr, float *loss_trace)
{
    if (!m || !gs || !gs->label || steps < 0) return RPBA_E_INVALID;
    LpGraphs G;
    int rc = lp_upload(m, gs, G);
    if (rc) return rc;
    for (int s = 0; s < steps; ++s) {
        m->model.forward_backward(G, m->st);
        m->model.adam(lr, m->st);
        if (loss_trace) {
            std::vector<float> l(G.ng);
            if (hipMemcpyAsync(l.data(), m->model.loss, sizeof(float) * G.ng, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
                hipStreamSynchronize(m->st) != hipSuccess)
                return RPBA_E_HIP;
            double t = 0.;
            for (float x : l) t += x;
            loss_trace[s] = (float)(t / G.ng);
        }
    }
    return hipStreamSynchronize(m->st) == hipSuccess ? RPBA_OK : RPBA_E_HIP;
}

// the model's prediction for the graph of the current S inside a solve
namespace rpba {
int lp_model_forward(rpba_lp *m, const LpGraphs &G, float *omega_dev, hipStream_t st)
{
    int rc = m->model.reserve(G.nn, 1);
    if (rc) return rc;
    m->model.forward(G, omega_dev, st);
    return 0;
}
}  // namespace rpba
