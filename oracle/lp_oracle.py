"""CPU restatement of the learned preconditioner's graph network, for the
gradient checks of the HIP kernels.  TEST INFRASTRUCTURE ONLY: imported by
tests/, never by the product package.

The network (one weight w per graph):
    h0 = relu(W_in x + b_in)
    3 x: m = (sum_e a_e h_nbr(e)) / sum_e a_e;  h <- h + relu(W_s h + W_n m + b)
    w = 0.95 sigmoid(v . mean(h) + c)
with the BF16 storage of the HIP kernels emulated: weights, features, the
pre-activations, m and every h are rounded to BF16 (round to nearest even);
the sums run in float64 here (FP32 on the GPU).  Gradients come from torch's
autograd in float64 with the roundings passed straight through, which is how
the HIP backward treats them.  Parameter layout as rpba_lp_n_params: W_in,
b_in, per layer (W_s, W_n, b), v, c.
"""
import numpy as np
import torch

F, H, L = 48, 64, 3
WMAX = 0.95
N_PARAMS = H * F + H + L * (2 * H * H + H) + H + 1


def bf16_round(a):
    """float32 -> nearest-even BF16, returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


class _Round(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return torch.from_numpy(bf16_round(x.detach().numpy().astype(np.float32)))

    @staticmethod
    def backward(ctx, g):
        return g


def split(p):
    p = torch.as_tensor(p, dtype=torch.float64)
    o = 0
    def take(n, shape):
        nonlocal o
        t = p[o:o + n].reshape(shape)
        o += n
        return t
    W_in = take(H * F, (H, F)); b_in = take(H, (H,))
    layers = []
    for _ in range(L):
        layers.append((take(H * H, (H, H)), take(H * H, (H, H)), take(H, (H,))))
    v = take(H, (H,)); c = take(1, (1,))
    return W_in, b_in, layers, v, c


def forward(params, graphs, rounding=True):
    """graphs: dict goff, feat [nn, F], adj_off, adj_nbr, adj_w (numpy).
    Returns omega (torch, per graph) with the graph of params (torch leaf)."""
    R = _Round.apply if rounding else (lambda t: t)
    W_in, b_in, layers, v, c = split(params)
    W_in, b_in, v, c = R(W_in), R(b_in), R(v), R(c)
    x = torch.from_numpy(np.asarray(graphs["feat"], dtype=np.float64))
    if rounding:
        x = torch.from_numpy(bf16_round(graphs["feat"]))
    nn = x.shape[0]
    off = np.asarray(graphs["adj_off"]); nbr = np.asarray(graphs["adj_nbr"]); w = np.asarray(graphs["adj_w"], np.float64)
    rows = np.repeat(np.arange(nn), np.diff(off))
    ns = np.bincount(rows, weights=w, minlength=nn)
    A = torch.zeros((nn, nn), dtype=torch.float64)
    if rows.size:
        A.index_put_((torch.from_numpy(rows), torch.from_numpy(nbr.astype(np.int64))),
                     torch.from_numpy(w), accumulate=True)
    scale = torch.from_numpy(np.where(ns > 0, 1.0 / np.where(ns > 0, ns, 1.0), 0.0))
    pre = R(x @ W_in.T + b_in)
    h = R(torch.relu(pre))
    for (Ws, Wn, b) in layers:
        Ws, Wn, b = R(Ws), R(Wn), R(b)
        m = R((A @ h) * scale[:, None])
        pre = R(h @ Ws.T + m @ Wn.T + b)
        h = R(h + torch.relu(pre))
    goff = np.asarray(graphs["goff"])
    out = []
    for g in range(goff.size - 1):
        gv = h[goff[g]:goff[g + 1]].mean(dim=0)
        out.append(WMAX * torch.sigmoid(gv @ v + c[0]))
    return torch.stack(out)


def loss_and_grad(params, graphs, labels, rounding=True):
    p = torch.tensor(np.asarray(params, dtype=np.float64), requires_grad=True)
    om = forward(p, graphs, rounding)
    y = torch.from_numpy(np.asarray(labels, dtype=np.float64))
    loss = ((om - y) ** 2).mean()
    loss.backward()
    return om.detach().numpy(), float(loss), p.grad.numpy()
