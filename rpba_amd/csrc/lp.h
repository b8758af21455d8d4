// Learned correction of the block-Jacobi preconditioner: a small graph
// network over the camera covisibility graph of the reduced camera system S
// predicts the weight w of the first-order Neumann correction
// M^-1 = C - w C N C (C = blockdiag(S_kk)^-1, N = S - blockdiag(S_kk)).
//
// Graph: one node per camera, one edge per non-zero off-diagonal block S_kl,
// weight e_kl = |S_kl|_F / sqrt(|S_kk|_F |S_ll|_F).  Node features (48): the
// nine log-diagonals of S_kk minus their mean, the 36 correlations
// S_kk[i][j] / sqrt(S_kk[i][i] S_kk[j][j]) (i < j), log(1 + degree) / 8,
// log(1 + sum_l e_kl) and max_l e_kl.
//
// Network (weights and stored activations BF16, arithmetic FP32):
//   h0 = relu(W_in x + b_in)                                  64
//   3 x: m = sum_l (e_kl / sum_l e_kl) h_l;  h <- h + relu(W_s h + W_n m + b)
//   g = mean over nodes of h;  w = 0.95 sigmoid(v . g + c)
// trained with Adam on (w - w*)^2, w* the weight with the fewest PCG
// iterations on a logged S.
#pragma once

#include <hip/hip_runtime.h>
#include "kernels.h"

#include <cstdint>
#include <vector>

namespace rpba {

constexpr int kLpF = 48;   // node features
constexpr int kLpH = 64;   // hidden width
constexpr int kLpL = 3;    // message-passing layers
constexpr float kLpWmax = 0.95f;
// parameter layout (floats): W_in [H][F], b_in [H], then per layer W_s [H][H],
// W_n [H][H], b [H], then v [H], c
constexpr int kLpParams = kLpH * kLpF + kLpH + kLpL * (2 * kLpH * kLpH + kLpH) + kLpH + 1;

// a set of graphs on the device, CSR adjacency over global node indices
struct LpGraphs {
    int ng = 0, nn = 0;
    long long ne = 0;
    const int *goff = nullptr;     // [ng + 1]
    const float *feat = nullptr;   // [nn][kLpF]
    const int *adj_off = nullptr;  // [nn + 1] (long ranges fit: edges < 2^31)
    const int *adj_nbr = nullptr;  // [ne]
    const float *adj_w = nullptr;  // [ne]
    const float *label = nullptr;  // [ng] or null
};

// features and CSR of the camera graph of S, from the block structure of the
// solve (upper blocks bk/bl, row_off, col_off/col_blk); adj_off[k] =
// row_off[k] + col_off[k]
void lp_graph_from_S(int nc, int nblk, const double *Sd, const double *Su, const SuMap &sm, const int *bk, const int *bl,
                     const int *row_off, const int *col_off, const int *col_blk, float *nrm, float *eb, float *feat,
                     int *adj_off, int *adj_nbr, float *adj_w, hipStream_t st);

struct LpModel {
    float *p = nullptr;          // FP32 master parameters
    uint16_t *pb = nullptr;      // BF16 copy used by the forward pass
    float *g = nullptr, *m = nullptr, *v = nullptr;  // gradient, Adam moments
    float *gpart = nullptr;      // per-chunk weight-gradient partials
    size_t gpart_cap = 0;
    // activations (BF16) and work arrays, sized for nn nodes / ng graphs
    uint16_t *act = nullptr;     // h0, (m_i, pre_i, h_{i+1}) per layer, pre_in
    float *dh = nullptr, *dpre = nullptr, *q = nullptr, *nsum = nullptr, *gvec = nullptr, *omega = nullptr;
    float *loss = nullptr;
    size_t cap_nodes = 0, cap_graphs = 0;
    int64_t adam_t = 0;
    ~LpModel();
    int init(const float *params, uint64_t seed);
    int reserve(int nn, int ng);
    // forward; omega_out[ng] on the device (may be model's own)
    void forward(const LpGraphs &G, float *omega_out, hipStream_t st);
    // forward + backward of sum_g (w_g - label_g)^2 / ng; gradient in g
    void forward_backward(const LpGraphs &G, hipStream_t st);
    void adam(float lr, hipStream_t st);
};

}  // namespace rpba

struct rpba_lp;
namespace rpba {
// the model's weight for the graph of one S, written to omega_dev on st
int lp_model_forward(rpba_lp *m, const LpGraphs &G, float *omega_dev, hipStream_t st);
}  // namespace rpba
