"""Pins of the oracle's partitioned solve (oracle/rpba_oracle_part.c) and the
product's image partitioner, on the CPU.

- A part's LSA step in GCP mode (rbacore.cpp:233-329, 1677-1729): the
  reduced camera system with the point priors equals a dense numpy build of
  the weighted normal equations (observation rows plus, per GCP point, three
  prior rows -I with weight W3 and residual centre - X) eliminated over the
  points; no gauge camera (every image 9 columns, image 0 included); the n-th
  kept GCP point takes entry n of the part's list; PointXXvIn's rule for a
  one-entry point (kept only with a GCP number > 0).
- CalcLSAXMulti (rpbacore.cpp:536-773): converges to the exact point on
  noise-free observations; the stored part information W3 equals the sum of
  B^T W B over the entries of the other parts (numpy derivatives by finite
  differences of the projection); the weight-only pass's shares in closed
  form; robust deletion of an entry past the threshold; a one-entry track is
  cleared with its point restored.
- The outer loop's control: the iflag and stopping sequence replayed from
  the oracle's own IT lines by the rule of rpbacore.cpp:479-493.
- The partitioner: disconnected camera groups come back as the parts; every
  part non-empty; deterministic; weight balance."""
import numpy as np
import pytest

from oracle import oracle
from rpba_amd import api, capi, scenes


def _proj(P, f, d, X):
    """CalcCorrPointMulti's prediction f z (u, v) (rpbacore.cpp:797-826)."""
    p = P[:, :3] @ X + P[:, 3]
    u, v = p[0] / p[2], p[1] / p[2]
    dist = u * u + v * v
    z = 1.0 + float(d[0]) * dist + float(d[1]) * dist * dist
    return np.array([float(f) * z * u, float(f) * z * v])


def _fd_jac(P, f, d, X, h=1e-6):
    J = np.zeros((2, 3))
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        J[:, c] = (_proj(P, f, d, X + e) - _proj(P, f, d, X - e)) / (2 * h)
    return J


def _wsym(w):
    return np.array([[w[0], w[3], w[4]], [w[3], w[1], w[5]], [w[4], w[5], w[2]]], dtype=np.float64)


# ------------------------------------------------------------------ blocks

def _tiny():
    return scenes.generate(None, seed=11, n_cams=6, n_pts=60, track="uniform2_8", outlier_frac=0.0)


@pytest.mark.parametrize("mult", [1e-4])
def test_block_step_normal_equations_with_priors(mult):
    sc = _tiny()
    n = sc["n_pts"]
    rng = np.random.default_rng(5)
    gcpin = -np.ones(n, dtype=np.int32)
    gpts = [3, 10, 20, 41]
    ids = [7, 2, 9, 4]  # global GCP numbers, any order
    for p, g in zip(gpts, ids):
        gcpin[p] = g
    gcpc = np.zeros((12, 3))
    X = sc["pts"].astype(np.float64)
    for p, g in zip(gpts, ids):
        gcpc[g] = X[p] + rng.normal(0, 0.05, 3)
    W3 = []
    for _ in gpts:
        A = rng.normal(0, 1, (3, 3))
        M = A @ A.T + 3 * np.eye(3)
        W3.append([M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2]])
    W3 = np.array(W3, dtype=np.float32)
    h = oracle.BlockHandle(sc, capi.make_flags(robust=0, k2=1, k4=1, f=1), gcpin, W3, gcpc)
    assert h.setup() == 0
    d = h.dims()
    cp = h.camparv()
    assert d["normcamera"] == -1
    assert np.all(np.diff(cp) == 9) and d["nrioff"] == 9 * sc["n_cams"]  # no gauge camera, image 0 free
    nrpv, gk = h.kept()
    assert np.array_equal(nrpv, np.arange(n))
    th0 = h.theta()
    r, _ = h.residuals(th0)
    Bc, Bp, ncol = h.jacobians()
    nc = d["nrioff"]
    J = np.zeros((2 * d["ntr"], d["nrioff3"]))
    # the oracle's tracks are point-major in file order: rebuild that order
    order = np.argsort(sc["obs_pt"], kind="stable")
    for q, fo in enumerate(order):
        k, p = sc["obs_cam"][fo], sc["obs_pt"][fo]
        J[2 * q:2 * q + 2, cp[k]:cp[k] + 9] = Bc[q, :, :9]
        J[2 * q:2 * q + 2, nc + 3 * p:nc + 3 * p + 3] = Bp[q]
    N = J.T @ J
    g = J.T @ r
    for m, p in enumerate(gpts):  # the m-th kept GCP point takes entry m
        Jp = np.zeros((3, d["nrioff3"]))
        Jp[:, nc + 3 * p:nc + 3 * p + 3] = -np.eye(3)
        Ws = _wsym(W3[m].astype(np.float64))
        N += Jp.T @ Ws @ Jp
        g += Jp.T @ Ws @ (gcpc[gcpin[p]] - X[p])
    Ucc, Ncp, Npp = N[:nc, :nc], N[:nc, nc:], N[nc:, nc:].copy()
    Npp[np.diag_indices_from(Npp)] *= 1.0 + mult
    Vinv = np.linalg.inv(Npp)
    S = Ucc - Ncp @ Vinv @ Ncp.T
    S[np.diag_indices_from(S)] *= 1.0 + mult
    S[np.diag_indices_from(S)] += 1e-2
    b = g[:nc] - Ncp @ Vinv @ g[nc:]
    So, bo = h.linearize(mult)
    assert np.abs(So - S).max() < 1e-9 * np.abs(S).max()
    assert np.abs(bo - b).max() < 1e-9 * np.abs(b).max()
    # a prior with the wrong weight entry would not match
    Ws_wrong = _wsym(W3[1].astype(np.float64))
    assert np.abs(Ws_wrong - _wsym(W3[0].astype(np.float64))).max() > 1e-3


def test_block_keep_rule_for_one_entry_points():
    """PointXXvIn in GCP mode keeps a point with a single entry only when its
    GCP number is > 0 (rbacore.cpp:1954-1956): number 0 needs two."""
    sc = _tiny()
    keep = np.ones(sc["n_obs"], dtype=bool)
    for p in (5, 6, 7):
        idx = np.nonzero(sc["obs_pt"] == p)[0]
        keep[idx[1:]] = False
    s2 = dict(sc, obs_cam=sc["obs_cam"][keep], obs_pt=sc["obs_pt"][keep], obs_xy=sc["obs_xy"][keep],
              n_obs=int(keep.sum()))
    gcpin = -np.ones(sc["n_pts"], dtype=np.int32)
    gcpin[5], gcpin[6] = 2, 0
    h = oracle.BlockHandle(s2, capi.make_flags(robust=0), gcpin, np.ones((4, 6), np.float32), np.zeros((4, 3)))
    nrpv, gk = h.kept()
    assert 5 in nrpv and 6 not in nrpv and 7 not in nrpv
    assert gk[list(nrpv).index(5)] == 2


# ------------------------------------------------------------ point kernel

def _point_setup(nerr=0):
    sc = _tiny()
    PM = sc["cam_P"].astype(np.float64).reshape(-1, 12)
    cf = sc["cam_f"].astype(np.float32)
    cd = sc["cam_d"].astype(np.float32).reshape(-1, 2)
    cams = np.array([0, 1, 2, 3, 4, 5], dtype=np.int32)
    Xt = sc["truth"]["X"][0].astype(np.float64)
    obs = np.array([_proj(PM[k].reshape(3, 4), cf[k], cd[k], Xt) for k in cams], dtype=np.float32)
    if nerr:
        obs[2, 1] += 30.0
    return PM, cf, cd, cams, obs, Xt


def test_lsax_converges_to_the_exact_point_and_stores_the_other_parts_information():
    PM, cf, cd, cams, obs, Xt = _point_setup()
    part = np.array([0, 0, 1, 1, 2, 2], dtype=np.int32)  # part of each image
    spart = np.array([0, 1, 2, 3], dtype=np.int32)       # slot 3: a part the point no longer sees
    pit = np.array([1, 1, 1, 0], dtype=np.uint8)
    X0 = Xt + np.array([0.05, -0.04, 0.03])
    conv, X, kept, w3, sg, rd = oracle.lsax_point(PM, cf, cd, obs, None, cams, np.arange(6), X0, part, spart, pit)
    assert conv == 1 and len(kept) == 6
    assert np.allclose(X, Xt, rtol=0, atol=1e-5 * np.abs(Xt).max())
    assert rd == 2 * 6 - 3 and sg < 1e-6
    # W3 of slot s: sum over the entries of the other parts of B^T W B (unit weights)
    Bs = [_fd_jac(PM[k].reshape(3, 4), cf[k], cd[k], X) for k in cams]
    for s in range(3):
        E = sum(B.T @ B for B, k in zip(Bs, cams) if part[k] != spart[s])
        assert np.allclose(_wsym(w3[s]), E, rtol=2e-3, atol=1e-6 * np.abs(E).max()), s
    assert np.all(w3[3] == 0)  # not seen: never written


def test_lsax_weight_only_pass_shares():
    PM, cf, cd, cams, obs, Xt = _point_setup()
    rng = np.random.default_rng(2)
    w = np.abs(rng.normal(1.0, 0.2, (6, 3))).astype(np.float32)
    w[:, 2] = 0.1 * w[:, 0]
    part = np.zeros(6, dtype=np.int32)
    X0 = Xt + 0.02
    conv, X, kept, w3, sg, rd = oracle.lsax_point(PM, cf, cd, obs, w, cams, np.arange(6), X0, part, [0], [1],
                                                  weightonly=1)
    assert np.array_equal(X, X0)  # weight-only: the point is not moved
    assert rd == 2 * 6 - 3
    r = np.array([obs[i].astype(np.float64) - _proj(PM[k].reshape(3, 4), cf[k], cd[k], X0) for i, k in enumerate(cams)])
    cost = sum(r[i, 0] ** 2 * w[i, 0] + r[i, 1] ** 2 * w[i, 1] + 2 * r[i, 0] * r[i, 1] * w[i, 2] for i in range(6))
    mean_w = (w[:, 0].astype(np.float64) + w[:, 1]).sum() / 12.0
    assert abs(sg - cost / mean_w) <= 1e-5 * cost / mean_w


def test_lsax_robust_deletion_and_clearing():
    PM, cf, cd, cams, obs, Xt = _point_setup(nerr=1)
    part = np.zeros(6, dtype=np.int32)
    X0 = Xt + 0.01
    conv, X, kept, w3, sg, rd = oracle.lsax_point(PM, cf, cd, obs, None, cams, np.arange(6), X0, part, [0], [1],
                                                  robustmode=2, sig0wv=np.full(6, 0.1, np.float32))
    assert conv == 1 and list(kept) == [0, 1, 3, 4, 5]
    assert rd == 2 * 5 - 3
    # a single entry: cleared, the point restored
    conv, X, kept, w3, sg, rd = oracle.lsax_point(PM, cf, cd, obs[:1], None, cams[:1], np.arange(1), X0, part, [0],
                                                  [1])
    assert conv == 0 and len(kept) == 0 and np.array_equal(X, X0.astype(np.float32).astype(np.float64))
    assert sg == 0 and rd == 0


# -------------------------------------------------------------- outer loop

@pytest.mark.parametrize("robust", [0, 1])
def test_outer_loop_control_replayed_from_its_trace(robust):
    sc = scenes.generate(None, seed=9, n_cams=150, n_pts=3000, track="uniform2_8", outlier_frac=0.02)
    part = api.partition_images(sc, 2)
    rc, out, st, log = oracle.partition_solve(sc, capi.make_flags(robust=robust, threads=2), part)
    assert rc == 0
    its = [l.split() for l in log if l.startswith("IT:")]
    assert len(its) == st["itt"]
    s = np.asarray(st["trace_sig0"][:st["itt"]])
    acc = np.asarray(st["trace_accept"][:st["itt"]])
    # the rule of rpbacore.cpp:242-245, 405-411, 479-493 on the oracle's sig0g
    iflag, sig0opt, frac = (2 if robust else 1), 1e20, float(np.float32(1.01))
    for k, t in enumerate(its):
        it, fl = int(t[1]), int(t[2])
        assert it == k + 1 and fl == iflag
        assert float(t[5]) == pytest.approx(s[k], rel=1e-5)
        if iflag == 1 and robust:
            frac = float(np.float32(1.02))
        assert acc[k] == (s[k] <= sig0opt)  # the state saved as the best
        if it > 1:
            if s[k] <= sig0opt:
                if sig0opt < s[k] * frac and it > 3:
                    iflag -= 1
                else:
                    sig0opt = s[k]
            elif it > 3:
                iflag -= 1
    assert iflag == 0 or len(its) == 50
    assert st["sig0_final"] == sig0opt


# ------------------------------------------------------------- partitioner

def _disjoint(groups=3, cams=40):
    parts = []
    for g in range(groups):
        parts.append(scenes.generate(None, seed=100 + g, n_cams=cams, n_pts=1500, track="uniform2_8", outlier_frac=0.0))
    cam, pt = [], []
    oc = op = 0
    for s in parts:
        cam.append(s["obs_cam"] + oc)
        pt.append(s["obs_pt"] + op)
        oc += s["n_cams"]
        op += s["n_pts"]
    return {"n_cams": oc, "n_pts": op, "obs_cam": np.concatenate(cam), "obs_pt": np.concatenate(pt)}


def test_partitioner_recovers_disconnected_groups():
    """Camera groups that share no point come back as the parts when there are
    as many parts as groups (a cut of zero edges at exact balance)."""
    for groups in (2, 3):
        sc = _disjoint(groups)
        part = api.partition_images(sc, groups)
        for g in range(groups):
            assert len(set(part[40 * g:40 * (g + 1)].tolist())) == 1
        assert len(set(part.tolist())) == groups


@pytest.mark.parametrize("nparts", [2, 3, 7])
def test_partitioner_valid_deterministic_balanced(nparts):
    sc = scenes.generate(None, seed=9, n_cams=150, n_pts=3000, track="uniform2_8", outlier_frac=0.0)
    a = api.partition_images(sc, nparts)
    b = api.partition_images(sc, nparts)
    assert np.array_equal(a, b)
    assert a.min() == 0 and a.max() == nparts - 1
    cnt = np.bincount(a, minlength=nparts)
    assert cnt.min() >= 1
    assert cnt.max() <= 1.25 * 150 / nparts + 2
