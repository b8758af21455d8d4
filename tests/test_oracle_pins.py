"""Pins of the CPU oracle against checks that do not restate its formulas:
finite differences, dense numpy linear algebra, scipy rotations, closed forms
and an independent replay of the published PCG variant.  CPU only."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle
from rpba_amd import capi, scenes


# --------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------

def test_cayley_identity_and_rotation_angle():
    # Cay(0) = I exactly (rbacore.cpp:2020-2040)
    assert np.array_equal(oracle.cayley(0.0, 0.0, 0.0), np.eye(3))
    rng = np.random.default_rng(0)
    for _ in range(20):
        v = rng.normal(size=3) * rng.uniform(0.01, 3.0)
        R = oracle.cayley(*v)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert np.isclose(np.linalg.det(R), 1.0, atol=1e-14)
        # axis is v, angle theta with tan(theta/2) = |v|/2
        assert np.allclose(R @ v, v, atol=1e-13 * max(1, np.linalg.norm(v)))
        theta = 2.0 * np.arctan(np.linalg.norm(v) / 2.0)
        assert np.isclose(np.trace(R), 1.0 + 2.0 * np.cos(theta), atol=1e-13)
        # direction: R = exp([v]x theta/|v|)
        Rs = Rotation.from_rotvec(v / np.linalg.norm(v) * theta).as_matrix()
        assert np.allclose(R, Rs, atol=1e-13)


def test_rectify_projects_onto_rotations():
    rng = np.random.default_rng(1)
    for _ in range(20):
        rv = Rotation.random(random_state=rng.integers(1 << 30)).as_rotvec()
        # the quaternion is built from the skew part, which vanishes at angle
        # pi, so the projection is only well conditioned away from pi
        rv *= min(1.0, 2.5 / np.linalg.norm(rv))
        R = Rotation.from_rotvec(rv).as_matrix()
        # exact rotations are (nearly) fixed points
        assert np.allclose(oracle.rectify(R), R, atol=1e-14)
        # a perturbed matrix goes back to an exact rotation close to it
        Rp = R + rng.normal(scale=1e-4, size=(3, 3))
        Rr = oracle.rectify(Rp)
        assert np.allclose(Rr @ Rr.T, np.eye(3), atol=1e-14)
        assert np.isclose(np.linalg.det(Rr), 1.0, atol=1e-14)
        assert np.abs(Rr - R).max() < 1e-3
        # the quaternion it builds is the one of the skew part: rectify is
        # idempotent up to rounding
        assert np.allclose(oracle.rectify(Rr), Rr, atol=1e-14)


def test_angle_axis_conversions_match_scipy():
    rng = np.random.default_rng(2)
    for _ in range(50):
        v = rng.normal(size=3)
        v *= rng.uniform(0.0, 3.0) / np.linalg.norm(v)  # angles below pi
        R = oracle.aa_to_R(*v)
        assert R.dtype == np.float32
        assert np.allclose(R, Rotation.from_rotvec(v).as_matrix(), atol=1e-6)
        back = oracle.R_to_aa(R)
        assert np.allclose(back, v, atol=2e-5)
    assert np.array_equal(oracle.aa_to_R(0.0, 0.0, 0.0), np.eye(3, dtype=np.float32))
    assert np.array_equal(oracle.R_to_aa(np.eye(3, dtype=np.float32)), np.zeros(3))
    # the near-pi branch (rpba.cpp:379-430)
    v = np.array([np.pi - 1e-3, 0.0, 0.0])
    assert np.allclose(np.abs(oracle.R_to_aa(oracle.aa_to_R(*v))), np.abs(v), atol=2e-2)


def test_small_inverses_match_numpy():
    rng = np.random.default_rng(3)
    for _ in range(20):
        A = rng.normal(size=(3, 3))
        A = A @ A.T + 0.1 * np.eye(3)
        assert np.allclose(oracle.inv3(A), np.linalg.inv(A), rtol=1e-12, atol=1e-12)
        for n in (3, 8, 9):
            B = rng.normal(size=(n, n)) + 0.5 * np.eye(n)
            assert np.allclose(oracle.invlu(B), np.linalg.inv(B), rtol=1e-9, atol=1e-10)
    # pivoting is needed: zero leading entry
    P = np.array([[0.0, 2.0, 1.0], [1.0, 1.0, 0.0], [3.0, 0.0, 1.0]])
    assert np.allclose(oracle.invlu(P), np.linalg.inv(P), atol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 10, 101])
def test_upper_median(n):
    rng = np.random.default_rng(n)
    v = rng.exponential(size=n).astype(np.float32)
    assert oracle.upper_median(v) == np.sort(v)[n // 2]


# --------------------------------------------------------------------------
# Jacobians by finite differences; Schur by dense normal equations
# --------------------------------------------------------------------------

def _probe(scene, k2=1, k4=1, f=1, solver=0):
    h = oracle.Handle(scene, capi.make_flags(robust=1, k2=k2, k4=k4, f=f, solver=solver))
    h.setup_pass0()
    return h


def _dense_jacobian(h, scene):
    d = h.dims()
    cp = h.camparv()
    Bc, Bp, ncol = h.jacobians()
    assert d["ntr"] == scene["n_obs"] and d["npts"] == scene["n_pts"]
    J = np.zeros((2 * d["ntr"], d["nrioff3"]))
    for q in range(d["ntr"]):
        k, p = scene["obs_cam"][q], scene["obs_pt"][q]
        n = ncol[q]
        assert n == cp[k + 1] - cp[k]
        J[2 * q:2 * q + 2, cp[k]:cp[k] + n] = Bc[q, :, :n]
        J[2 * q:2 * q + 2, d["nrioff"] + 3 * p:d["nrioff"] + 3 * p + 3] = Bp[q]
    return J


@pytest.mark.parametrize("flags", [(1, 1, 1), (1, 0, 1), (0, 0, 0), (0, 1, 0)])
def test_jacobian_matches_finite_differences(tiny_scene, flags):
    h = _probe(tiny_scene, *flags)
    d = h.dims()
    assert d["nraddpi"] == sum(flags)
    th0 = h.theta()
    r0, _ = h.residuals(th0)
    J = _dense_jacobian(h, tiny_scene)
    Jfd = np.zeros_like(J)
    for j in range(d["nrioff3"]):
        step = 1e-6 * max(1.0, abs(th0[j]))
        tp = th0.copy(); tp[j] += step
        tm = th0.copy(); tm[j] -= step
        Jfd[:, j] = (h.residuals(tp)[0] - h.residuals(tm)[0]) / (2 * step)
    h.residuals(th0)
    scale = np.abs(Jfd).max(axis=0, keepdims=True) + 1e-3
    err = np.abs(J - Jfd) / scale
    assert err.max() < 1e-5, (np.unravel_index(err.argmax(), err.shape), err.max())
    # the residual is b - p: a zero step gives the stored residuals
    assert np.allclose(r0, h.residuals(th0)[0])


def test_gauge_columns(tiny_scene):
    """Image 0 carries only the intrinsics; the gauge camera loses the C
    component of largest magnitude (rbacore.cpp:1042-1069)."""
    h = _probe(tiny_scene)
    d = h.dims()
    cp = h.camparv()
    sizes = np.diff(cp)
    assert sizes[0] == d["nraddpi"]
    nc = d["normcamera"]
    assert sizes[nc] == 8 and all(sizes[k] == 9 for k in range(1, d["nimg"]) if k != nc)
    assert d["nrioff"] == d["nraddpi"] + (5 + d["nraddpi"]) + (6 + d["nraddpi"]) * (d["nimg"] - 2)
    # gauge camera: the farthest from image 0 by |R_i t_i - R_0 t_0|
    P = tiny_scene["cam_P"].astype(np.float64)
    T = np.einsum("nij,nj->ni", P[:, :, :3], P[:, :, 3])
    dist = ((T[1:] - T[0]) ** 2).sum(axis=1)
    assert nc == 1 + int(np.argmax(dist))


@pytest.mark.parametrize("mult", [1e-3, 1e-6])
def test_schur_matches_dense_elimination(tiny_scene, mult):
    h = _probe(tiny_scene)
    d = h.dims()
    th0 = h.theta()
    r, _ = h.residuals(th0)
    J = _dense_jacobian(h, tiny_scene)
    nc = d["nrioff"]
    N = J.T @ J           # unit weights
    g = J.T @ r
    Ucc, Ncp, Npp = N[:nc, :nc], N[:nc, nc:], N[nc:, nc:].copy()
    # point blocks are 3x3 block diagonal; damp their diagonals
    Npp[np.diag_indices_from(Npp)] *= 1.0 + mult
    Vinv = np.linalg.inv(Npp)
    S = Ucc - Ncp @ Vinv @ Ncp.T
    S[np.diag_indices_from(S)] *= 1.0 + mult
    S[np.diag_indices_from(S)] += 1e-2
    b = g[:nc] - Ncp @ Vinv @ g[nc:]
    So, bo = h.linearize(mult)
    assert np.allclose(So, So.T, rtol=0, atol=1e-9 * np.abs(So).max())
    assert np.abs(So - S).max() < 1e-9 * np.abs(S).max()
    assert np.abs(bo - b).max() < 1e-9 * np.abs(b).max()

    # one full step: camera part by PCG (loose), point part by back-substitution
    x, its = h.step(mult)
    assert its >= 1
    # the reference's step-norm rule stops loosely (a ~316x smaller step), so
    # the in-solve PCG is checked against the replay on the same S and b
    cp = h.camparv()
    xr, _ = _replay_pcg(np.diff(cp), So, bo)
    assert np.allclose(x[:nc], xr.astype(np.float32), rtol=1e-5, atol=1e-6 * np.abs(xr).max())
    assert np.linalg.norm(S @ x[:nc] - b) < np.linalg.norm(b)
    # camera step is cast to float (rbacore.cpp:1185)
    assert np.array_equal(x[:nc], x[:nc].astype(np.float32).astype(np.float64))
    xp = Vinv @ (g[nc:] - Ncp.T @ x[:nc])
    assert np.allclose(x[nc:], xp, rtol=1e-4, atol=1e-6 * np.abs(xp).max())


def _weighted_tiny(img_size=None):
    return scenes.generate(None, seed=13, n_cams=6, n_pts=60, track="uniform2_8", outlier_frac=0.0,
                           weights="random", img_size=img_size)


def _weight_blocks(scene):
    """Block-diagonal observation weight matrix [[wx, wxy], [wxy, wy]] in the
    file order of the observations (PrepareWeights2D, rbacore.cpp:704-783)."""
    w = scene["obs_w"].astype(np.float64)
    n = w.shape[0]
    Wm = np.zeros((2 * n, 2 * n))
    for q in range(n):
        Wm[2 * q:2 * q + 2, 2 * q:2 * q + 2] = [[w[q, 0], w[q, 2]], [w[q, 2], w[q, 1]]]
    return Wm


def test_weighted_normal_equations_match_dense_elimination():
    """Non-unit weights with a cross term: U/V/W/g and the Schur step equal
    the dense elimination of J^T W J (the wxy term enters Bc^T W, Bp^T W and
    both off-diagonal entries, rbacore.cpp:1540-1609)."""
    sc = _weighted_tiny()
    h = _probe(sc)
    d = h.dims()
    r, _ = h.residuals(h.theta())
    J = _dense_jacobian(h, sc)
    Wm = _weight_blocks(sc)
    nc = d["nrioff"]
    mult = 1e-3
    N = J.T @ Wm @ J
    g = J.T @ Wm @ r
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
    # the same scene with the cross term dropped gives another S: the term matters
    sc0 = dict(sc, obs_w=sc["obs_w"] * np.array([1, 1, 0], np.float32))
    S0, _ = _probe(sc0).linearize(mult)
    assert np.abs(S0 - So).max() > 1e-4 * np.abs(So).max()


def test_weighted_keys_and_residual_clamp():
    """The robust key of an observation is r^T W r under its original
    weights (sig0rmv, rbacore.cpp:1312-1327), and residual components are
    clamped at 5000 / sqrt(ihw2) with ihw2 the mean of ((w + h) / 2)^2 over the
    images (rbacore.cpp:48, 1273-1289; rpbacore.cpp:87-90)."""
    sc = _weighted_tiny(img_size=(640, 480))
    xy = sc["obs_xy"].copy()
    xy[:5, 0] += 60.0  # far beyond the clamp
    xy[5:10, 1] -= 60.0
    sc = dict(sc, obs_xy=xy)
    h = _probe(sc)
    r, key = h.residuals(h.theta())
    r = r.reshape(-1, 2)
    w = sc["obs_w"].astype(np.float64)
    quad = r[:, 0] ** 2 * w[:, 0] + 2 * r[:, 0] * r[:, 1] * w[:, 2] + r[:, 1] ** 2 * w[:, 1]
    # keys are float: the float rounding of the residuals and products only
    assert np.allclose(key, quad, rtol=2e-6, atol=0)
    ihw2 = np.float32(0)
    for wi, hi in zip(sc["img_w"], sc["img_h"]):
        ihw2 += np.float32((int(hi) + int(wi)) ** 2) * np.float32(0.25)
    ihw2 /= np.float32(sc["n_cams"])
    clamp = 5000.0 / float(np.sqrt(ihw2))
    assert 8.0 < clamp < 9.5
    assert np.all(np.abs(r) <= clamp)
    assert np.all(r[:5, 0] == clamp) and np.all(r[5:10, 1] == -clamp)


# --------------------------------------------------------------------------
# PCG: against dense Cholesky and an independent replay of the rule
# --------------------------------------------------------------------------

def _replay_pcg(bs, S, b, omega=0.0):
    """Block-Jacobi PCG, Polak-Ribiere beta clipped at 0, stop when the step
    norm |alpha A d|^2 falls 1e5 below the first one, return the iterate with
    the smallest step norm; at most max(1000, n) iterations.  omega > 0: the
    preconditioner C - omega C N C with N the off-diagonal blocks of S."""
    off = np.concatenate([[0], np.cumsum(bs)])
    C = np.zeros_like(S)
    for i in range(len(bs)):
        sl = slice(off[i], off[i + 1])
        C[sl, sl] = np.linalg.inv(S[sl, sl])
    if omega:
        N = S.copy()
        for i in range(len(bs)):
            sl = slice(off[i], off[i + 1])
            N[sl, sl] = 0.0
        C = C - omega * C @ N @ C
    x = np.zeros_like(b)
    r = b.copy()
    h = C @ r
    d = h.copy()
    best, mbest, m0 = x.copy(), 1e50, None
    for k in range(max(1000, len(b))):
        Ad = S @ d
        rh = r @ h
        alpha = 0.0 if (abs(rh) < 1e-20 or abs(rh) > 1e20) else rh / (d @ Ad)
        x = x + alpha * d
        dr = alpha * Ad
        r1 = r - dr
        m = dr @ dr
        h1 = C @ r1
        if k == 0:
            m0 = m
        if m < mbest:
            best, mbest = x.copy(), m
        if m * 1e5 < m0:
            return best, k + 1
        num = r1 @ (h1 - h)
        beta = 0.0 if abs(num) < 1e-10 else num / rh
        beta = max(beta, 0.0)
        d = h1 + beta * d
        r, h = r1, h1
    return best, k + 1


@pytest.mark.parametrize("seed", range(6))
def test_pcg_against_cholesky_and_replay(seed):
    rng = np.random.default_rng(100 + seed)
    bs = np.array([3, 8] + [9] * int(rng.integers(3, 12)))
    n = bs.sum()
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    # moderate conditioning keeps rounding from moving the stopping iteration
    ev = np.logspace(0, rng.uniform(1, 2.5), n)
    S = (Q * ev) @ Q.T
    S = 0.5 * (S + S.T)
    b = rng.normal(size=n)
    x, its = oracle.pcg_dense(bs, S, b)
    xr, itr = _replay_pcg(bs, S, b)
    assert its == itr
    assert np.allclose(x, xr, rtol=1e-4, atol=1e-6 * np.abs(xr).max())
    xs = np.linalg.solve(S, b)
    # the step-norm rule stops loosely: residual reduced well below b
    assert np.linalg.norm(S @ x - b) < 0.1 * np.linalg.norm(b)
    assert np.linalg.norm(x - xs) < 0.2 * np.linalg.norm(xs)


@pytest.mark.parametrize("seed,omega", [(0, 0.5), (1, 0.75), (2, 0.3)])
def test_neumann_pcg_against_replay(seed, omega):
    """The corrected preconditioner (preconditioner = 1): same iteration count
    and iterate as the numpy replay with M^-1 = C - omega C N C, and fewer
    iterations than block-Jacobi on a block-coupled SPD system."""
    rng = np.random.default_rng(200 + seed)
    bs = np.array([9] * int(rng.integers(6, 14)))
    n = bs.sum()
    A = rng.normal(size=(n, 3 * n))
    S = A @ A.T / n + 0.05 * np.eye(n)
    b = rng.normal(size=n)
    x, its = oracle.pcg_dense(bs, S, b, omega)
    xr, itr = _replay_pcg(bs, S, b, omega)
    assert its == itr
    assert np.allclose(x, xr, rtol=1e-6, atol=1e-8 * np.abs(xr).max())
    x0, it0 = oracle.pcg_dense(bs, S, b)
    assert np.linalg.norm(S @ x - b) < 0.1 * np.linalg.norm(b)


@pytest.mark.parametrize("seed", range(4))
def test_dense_cholesky_against_numpy(seed):
    """The dense form's solve: factor equal to numpy's Cholesky, solution to
    numpy's solve; an indefinite matrix is refused."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(5, 80))
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    S = (Q * np.logspace(-2, 4, n)) @ Q.T
    S = 0.5 * (S + S.T)
    b = rng.normal(size=n)
    x, L = oracle.chol_solve(S, b)
    Ln = np.linalg.cholesky(S)
    assert np.abs(L - Ln).max() < 1e-10 * np.abs(Ln).max()
    xs = np.linalg.solve(S, b)
    assert np.abs(x - xs).max() < 1e-8 * np.abs(xs).max()
    S2 = S.copy()
    S2[n // 2, n // 2] = -1.0
    assert oracle.chol_solve(S2, b) is None


def test_dense_form_step_is_exact(tiny_scene):
    """solver = 1 replaces the PCG by the exact solve of the same S and b:
    the camera step equals numpy's solve (float-cast, H3) and the point step
    follows by back-substitution."""
    h = _probe(tiny_scene, solver=1)
    d = h.dims()
    nc = d["nrioff"]
    So, bo = h.linearize(1e-3)
    x, its = h.step(1e-3)
    assert its == 0
    xs = np.linalg.solve(So, bo)
    # float cast of the camera step (rbacore.cpp:1185): within one float ulp
    assert np.array_equal(x[:nc], x[:nc].astype(np.float32).astype(np.float64))
    assert np.abs(x[:nc] - xs).max() < 1e-6 * np.abs(xs).max()


def test_dense_form_solve_converges(scene_c1):
    """Whole robust solve with the exact camera step: a final sigma0 within
    1e-3 of the PCG solve's and the same robust labels up to the borderline
    observations of a different trajectory (at most 0.1%)."""
    f0 = capi.make_flags(robust=1)
    f1 = capi.make_flags(robust=1, solver=1)
    rc0, out0, st0, _ = oracle.solve(scene_c1, f0)
    rc1, out1, st1, _ = oracle.solve(scene_c1, f1)
    assert rc0 == 0 and rc1 == 0
    assert st1["pcg_iters"] == 0
    assert abs(st1["sigma_out"][1] - st0["sigma_out"][1]) < 1e-3 * st0["sigma_out"][1]
    assert (out0["obs_kept"] != out1["obs_kept"]).sum() <= 1e-3 * out0["obs_kept"].size


# --------------------------------------------------------------------------
# robust step and whole-solve invariants
# --------------------------------------------------------------------------

def test_robust_labels_and_passes(scene_c1):
    rc, out, st, log = oracle.solve(scene_c1, capi.make_flags(robust=1))
    assert rc == 0 and st["passes_done"] == 2
    n = scene_c1["n_obs"]
    # first iteration runs with iflag = 2: the 0.001 factor keeps every +-50 px
    # outlier under the kf = 9 threshold; the second flags them
    nr = st["trace_nrobs"]
    assert nr[0] == 2 * n
    first = int(np.argmax(nr < 2 * n))
    assert first >= 1 and nr[first] < 2 * n
    # the drop happens once progress stalls (iflag 2 -> 1): sig0 fell by less
    # than the 1% sig0frac between the two preceding trials
    s = st["trace_sig0"]
    assert s[first - 2 if first >= 2 else 0] < s[first - 1] * 1.01 + 1e-12 or first == 1
    truth = scene_c1["truth"]["outlier"]
    kept = out["obs_kept"].astype(bool)
    assert (~kept & truth).sum() >= 0.95 * truth.sum()
    # RobustDelete's rule seen on the output (the rule itself is pinned on
    # constructed labels in test_oracle_control.py): a kept point that lost
    # observations keeps at least minimagenrdel = 3 and a dropped point keeps
    # none
    t_all = np.bincount(scene_c1["obs_pt"], minlength=scene_c1["n_pts"])
    t_kept = np.bincount(scene_c1["obs_pt"], weights=kept, minlength=scene_c1["n_pts"]).astype(int)
    pk = out["pt_kept"].astype(bool)
    lost = t_kept < t_all
    assert (t_kept[pk & lost] >= 3).all()
    assert (t_kept[~pk] == 0).all() and (~pk).any() and (pk & lost).any()
    # converged to the noise level (0.5 px)
    assert abs(st["sigma_out"][1] - 0.5) < 0.02
    # log lines in the reference's format (rbacore.cpp:586-597)
    assert log[0].startswith("Sigma0: ") and "IT: 0  ITT: " in log[0]
    assert log[1].startswith("Sigma0: 1e+10 ")


def test_plain_vs_float_sites_agree(scene_c1):
    _, o0, s0, _ = oracle.solve(scene_c1, capi.make_flags(robust=1), fsites=0)
    _, o1, s1, _ = oracle.solve(scene_c1, capi.make_flags(robust=1), fsites=1)
    assert s0["itt"] == s1["itt"]
    assert np.array_equal(o0["obs_kept"], o1["obs_kept"])
    assert np.allclose(s0["trace_sig0"], s1["trace_sig0"], rtol=1e-5)


def test_zero_noise_scene_has_zero_residual():
    sc = scenes.generate(None, seed=7, n_cams=8, n_pts=200, track="uniform2_8", outlier_frac=0.0,
                         noise_px=0.0, rot_noise=0.0, t_noise=0.0, x_noise=0.0)
    # exact values up to the float quantisation of the inputs
    s, nobs, dof = oracle.exteval(sc, capi.make_flags(exteval=1, k2=1, k4=1, f=1))
    assert s < 1e-3
    assert nobs == 2 * sc["n_obs"]


def test_exteval_dof_formula(scene_c1):
    s, nobs, dof = oracle.exteval(scene_c1, capi.make_flags(exteval=1, k2=1, k4=1, f=1))
    t = np.bincount(scene_c1["obs_pt"])
    t = t[t > 1]
    # rpbacore.cpp:931-936
    assert nobs == 2 * t.sum()
    assert dof == (2 * t - 3).sum() - 6 * (50 - 2) - 5 - 50 * 3
    # sigma0 of the initial values is far above the 0.5 px noise
    assert s > 2.0


def test_error_paths():
    sc = scenes.generate(None, seed=8, n_cams=6, n_pts=80, track="uniform2_8", outlier_frac=0.0)
    # an image without observations (rbacore.cpp:414-425)
    m = sc["obs_cam"] != 5
    bad = dict(sc, obs_cam=sc["obs_cam"][m], obs_pt=sc["obs_pt"][m], obs_xy=sc["obs_xy"][m], n_obs=int(m.sum()))
    rc, out, st, log = oracle.solve(bad, capi.make_flags(robust=0))
    assert rc == capi.RPBA_E_NO_OBS_IN_IMAGE and "No observations in image 5" in "".join(log)
    # outputs still written from the (initial) stored state
    assert np.allclose(out["pts"][out["pt_kept"] == 1], sc["pts"][out["pt_kept"] == 1], atol=1e-6)
    # too few observations for the unknowns (rbacore.cpp:468-476)
    small = scenes.generate(None, seed=9, n_cams=6, n_pts=12, track="uniform2_8", outlier_frac=0.0)
    small["obs_cam"] = small["obs_cam"]
    rc, out, st, log = oracle.solve(small, capi.make_flags(robust=0))
    if 2 * small["n_obs"] < 3 * 12 + 3 + 8 + 9 * 4 + 3 * 6:
        assert rc == capi.RPBA_E_LOW_REDUNDANCY and "Low redundancy" in "".join(log)
