"""Learned preconditioner on the GPU: the network's forward and gradient
against the CPU oracle (torch autograd in float64 with the same BF16
roundings); training reproducible and reducing the loss; the corrected PCG
(fixed Neumann weight) against the oracle's solve; and what the trained
network gains over block-Jacobi on held-out scenes, with the PCG stopping
rule unchanged."""
import os

import numpy as np
import pytest

from oracle import lp_oracle as lo
from oracle import oracle
from rpba_amd import api, capi, lp_train, scenes
from test_gpu_parity import _cmp_solves
from test_lp_oracle import _graph

pytestmark = pytest.mark.gpu


def _as_list(g):
    out = []
    goff = g["goff"]
    for i in range(len(goff) - 1):
        n0, n1 = goff[i], goff[i + 1]
        e0, e1 = g["adj_off"][n0], g["adj_off"][n1]
        out.append({"feat": g["feat"][n0:n1], "adj_off": np.asarray(g["adj_off"][n0:n1 + 1]) - e0,
                    "adj_nbr": np.asarray(g["adj_nbr"][e0:e1]) - n0, "adj_w": g["adj_w"][e0:e1]})
    return out


def test_forward_and_gradient_match_oracle():
    rng = np.random.default_rng(5)
    g = _graph(rng, [6, 17, 40])
    y = np.array([0.2, 0.5, 0.85], dtype=np.float32)
    m = api.LpModel(seed=3)
    p = m.params()
    om, loss, grad = m.gradient(_as_list(g), y)
    om0, loss0, grad0 = lo.loss_and_grad(p, g, y, rounding=True)
    print("[lp grad] omega gpu", om, "cpu", om0, flush=True)
    assert np.abs(om - om0).max() < 1e-6
    assert abs(loss - loss0) < 1e-6 * max(loss0, 1e-3)
    rel = np.linalg.norm(grad - grad0) / np.linalg.norm(grad0)
    print(f"[lp grad] |g_gpu - g_cpu| / |g_cpu| = {rel:.2e}", flush=True)
    assert rel < 1e-5  # measured 5.7e-7
    # every parameter block on its own (input layer, each message-passing
    # layer's W_s, W_n and b, the readout), so that an error confined to one
    # layer's backward cannot hide in the norm of the whole gradient
    F, H, L = lo.F, lo.H, lo.L
    sizes = [("W_in", H * F), ("b_in", H)]
    for l in range(L):
        sizes += [(f"W_s{l}", H * H), (f"W_n{l}", H * H), (f"b{l}", H)]
    sizes += [("v", H), ("c", 1)]
    o = 0
    worst = 0.0
    for name, n in sizes:
        a, b = grad[o:o + n], grad0[o:o + n]
        o += n
        rb = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        worst = max(worst, rb)
        print(f"[lp grad] block {name}: rel {rb:.2e}", flush=True)
        assert rb < 1e-5, name  # measured at most 7.4e-7
    assert o == grad.size
    print(f"[lp grad] worst block {worst:.2e}", flush=True)


def test_training_reproducible_and_descends():
    rng = np.random.default_rng(9)
    g = _as_list(_graph(rng, [8, 12, 20, 30]))
    y = np.array([0.6, 0.4, 0.7, 0.5], dtype=np.float32)
    a = api.LpModel(seed=1)
    ta = a.train(g, y, 60, 3e-3)
    b = api.LpModel(seed=1)
    tb = b.train(g, y, 60, 3e-3)
    assert np.array_equal(a.params(), b.params())
    assert ta[-1] < 0.5 * ta[0]


def test_corrected_pcg_matches_oracle(scene_c1):
    _cmp_solves(scene_c1, capi.make_flags(robust=1, preconditioner=1, pc_omega=0.6), "c1 neumann 0.6")


def test_learned_preconditioner_gain_on_heldout():
    if not os.path.exists(lp_train.WEIGHTS):
        pytest.skip("no trained weights (run python -m rpba_amd.lp_train on a GPU)")
    model = api.LpModel.load(lp_train.WEIGHTS)
    ev = lp_train.evaluate(model, range(lp_train.HOLDOUT_SEED0, lp_train.HOLDOUT_SEED0 + 4))
    print("[lp gain] held-out PCG iterations", ev, flush=True)
    assert ev["learned"] < ev["block_jacobi"]
    # whole solves with the network in the loop: same converged sigma0
    sc = lp_train.scene_for(lp_train.HOLDOUT_SEED0 + 5)
    c0 = api.Context(sc, capi.make_flags(robust=1))
    c0.run(-1)
    _, _, s0 = c0.download()
    c2 = api.Context(sc, capi.make_flags(robust=1, preconditioner=2))
    c2.set_lp(model)
    c2.run(-1)
    _, _, s2 = c2.download()
    print(f"[lp gain] solve: PCG {s0['pcg_iters']} -> {s2['pcg_iters']}, sigma0 {s0['sigma_out'][1]:.6f} "
          f"-> {s2['sigma_out'][1]:.6f}, last w {c2.last_omega():.3f}", flush=True)
    assert abs(s2["sigma_out"][1] - s0["sigma_out"][1]) < 1e-3 * s0["sigma_out"][1]
