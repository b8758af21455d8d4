"""The CPU restatement of the learned preconditioner's network: closed form
of the readout, autograd against finite differences (float64, no BF16
rounding) and the BF16 rounding rule.  CPU only."""
import numpy as np

from oracle import lp_oracle as lo


def _graph(rng, sizes):
    feats, offs, nbrs, ws = [], [0], [], []
    goff = [0]
    base = 0
    for n in sizes:
        A = np.triu(rng.uniform(0, 1, (n, n)) * (rng.uniform(0, 1, (n, n)) < 0.5), 1)
        A = A + A.T
        for k in range(n):
            nz = np.nonzero(A[k])[0]
            nbrs.extend((nz + base).tolist())
            ws.extend(A[k, nz].tolist())
            offs.append(len(nbrs))
        feats.append(rng.normal(size=(n, lo.F)))
        base += n
        goff.append(base)
    return {"goff": np.array(goff), "feat": np.concatenate(feats).astype(np.float32), "adj_off": np.array(offs),
            "adj_nbr": np.array(nbrs, dtype=np.int64), "adj_w": np.array(ws, dtype=np.float32)}


def test_readout_closed_form():
    """All weights zero except c: every graph gets w = 0.95 sigmoid(c)."""
    rng = np.random.default_rng(0)
    g = _graph(rng, [4, 7])
    p = np.zeros(lo.N_PARAMS)
    p[-1] = 0.3
    om = lo.forward(p, g, rounding=False).detach().numpy()
    assert np.allclose(om, 0.95 / (1 + np.exp(-0.3)), rtol=0, atol=1e-15)


def test_autograd_against_finite_differences():
    rng = np.random.default_rng(1)
    g = _graph(rng, [5, 9, 3])
    p = rng.normal(size=lo.N_PARAMS) * 0.2
    y = np.array([0.3, 0.6, 0.8])
    _, loss, grad = lo.loss_and_grad(p, g, y, rounding=False)
    idx = rng.choice(lo.N_PARAMS, 40, replace=False)
    idx = np.concatenate([idx, [lo.N_PARAMS - 1, lo.N_PARAMS - 2, 0]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        lp = lo.loss_and_grad(p + e, g, y, rounding=False)[1]
        lm = lo.loss_and_grad(p - e, g, y, rounding=False)[1]
        fd = (lp - lm) / 2e-6
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, grad[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.14159, 0.0, 1.0 + 3 * 2 ** -8], dtype=np.float32)
    r = lo.bf16_round(x)
    # 8 significant bits (step 2^-7 at 1); ties to even
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[4] == 0.0 and r[5] == 1.0 + 2 ** -6
    assert abs(r[3] + 3.140625) < 1e-12
