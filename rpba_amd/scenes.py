"""Seeded synthetic bundle-adjustment scenes in the reference's data model.

The reference ships no data (its `parameters` file points at the BAL problem
problem-961-187103-pre.txt, which is not in the repository), so the workloads
are generated.  Scenes follow the BAL camera model as InputBlock reads it
(rpba.cpp:180-284): per image an angle-axis rotation (double), a translation,
f, k1, k2 (floats); observations (camera, point, x, y) in file order; points
XYZ.  With ``k24out = 1`` the reference uses k1, k2 unscaled
(rpba.cpp:255-258), which is what the generator writes.

Geometry: cameras on a circle of radius 10 in the z = 0 plane looking at the
origin, points uniform in a ball of radius 3.  Camera frame as in BAL:
P = R X + t, visible points have P_z < 0, and the reference's residual model is
xg = (-x_file, -y_file) = f s (P_x/P_z, P_y/P_z) with s = 1 + k1 rho + k2 rho^2
(rbacore.cpp:1238-1271).
"""
import numpy as np

# Workloads, one seed each.  Track lengths: uniform 2..8 for the 50-camera
# scene, else 2 + geometric with the stated mean.
CONFIGS = {
    "c1_50": dict(seed=1, n_cams=50, n_pts=5000, track="uniform2_8", outlier_frac=0.02),
    "c2_961": dict(seed=2, n_cams=961, n_pts=187103, track_mean=9.0, outlier_frac=0.03),
    "c3_4585": dict(seed=3, n_cams=4585, n_pts=1324582, track_mean=7.0, outlier_frac=0.03),
    "c5_13682": dict(seed=5, n_cams=13682, n_pts=4456117, track_mean=6.5, outlier_frac=0.03),
}


# Batched blocks: 64 independent blocks, cameras per block log-uniform in
# [20, 400], points 150 per camera clipped to [2000, 60000], track length
# 2 + geometric with mean 6.  Block b has seed 4000 + b.
BATCH_CONFIG = dict(seed=4, n_blocks=64, cam_lo=20, cam_hi=400, pts_per_cam=150, pts_lo=2000, pts_hi=60000,
                    track_mean=6.0, outlier_frac=0.02)


def batch_block(b, seed=4, n_blocks=64, cam_lo=20, cam_hi=400, pts_per_cam=150, pts_lo=2000, pts_hi=60000,
                track_mean=6.0, outlier_frac=0.02):
    """Block b of the batched-blocks workload alone."""
    rng = np.random.default_rng(seed)
    nc = int(np.round(np.exp(rng.uniform(np.log(cam_lo), np.log(cam_hi), n_blocks)))[b])
    npt = int(np.clip(pts_per_cam * nc, pts_lo, pts_hi))
    return generate(None, seed=1000 * seed + b, n_cams=nc, n_pts=npt, track="geometric",
                    track_mean=track_mean, outlier_frac=outlier_frac)


def generate_batch(seed=4, n_blocks=64, **kw):
    """Independent blocks of uneven size (the batched-blocks workload)."""
    return [batch_block(b, seed=seed, n_blocks=n_blocks, **kw) for b in range(n_blocks)]


def rodrigues(aa):
    """Angle-axis (n,3) -> rotation matrices (n,3,3), double, with the
    formula InputBlock uses (rpba.cpp:343-366)."""
    aa = np.asarray(aa, dtype=np.float64).reshape(-1, 3)
    a, b, c = aa[:, 0], aa[:, 1], aa[:, 2]
    a2, b2, c2 = a * a, b * b, c * c
    th = np.sqrt(a2 + b2 + c2)
    safe = np.where(th == 0.0, 1.0, th)
    ct = np.where(th == 0.0, 0.5, (1.0 - np.cos(th)) / (safe * safe))
    st = np.where(th == 0.0, 1.0, np.sin(th) / safe)
    R = np.empty((aa.shape[0], 3, 3))
    R[:, 0, 0] = 1.0 - (b2 + c2) * ct
    R[:, 0, 1] = a * b * ct - c * st
    R[:, 0, 2] = a * c * ct + b * st
    R[:, 1, 0] = a * b * ct + c * st
    R[:, 1, 1] = 1.0 - (c2 + a2) * ct
    R[:, 1, 2] = b * c * ct - a * st
    R[:, 2, 0] = a * c * ct - b * st
    R[:, 2, 1] = b * c * ct + a * st
    R[:, 2, 2] = 1.0 - (a2 + b2) * ct
    return R


def rot_to_aa(R):
    """Rotation matrices -> angle-axis, robust at every angle (quaternion
    based, via scipy)."""
    from scipy.spatial.transform import Rotation
    R = np.asarray(R, dtype=np.float64).reshape(-1, 3, 3)
    return Rotation.from_matrix(R).as_rotvec()


def _look_at_origin(centers):
    z = centers / np.linalg.norm(centers, axis=1, keepdims=True)  # camera +z axis points away from the scene
    up = np.array([0.0, 0.0, 1.0])
    x = np.cross(up[None, :], z)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(z, x)
    return np.stack([x, y, z], axis=1)  # rows = camera axes in world coordinates


def _window_tracks(rng, n_pts, n_cams, lengths, window):
    """Distinct cameras per point from a window of `window` consecutive
    cameras (on the circle) around a random first camera, sorted: banded
    covisibility, so the reduced camera system is sparse."""
    cams = np.empty(int(lengths.sum()), dtype=np.int32)
    off = np.zeros(n_pts + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    start = rng.integers(0, n_cams, n_pts)
    for t in np.unique(lengths):
        idx = np.nonzero(lengths == t)[0]
        pick = np.argsort(rng.random((idx.size, window)), axis=1)[:, :t]
        rows = np.sort((start[idx][:, None] + pick) % n_cams, axis=1)
        pos = off[idx][:, None] + np.arange(t)[None, :]
        cams[pos.ravel()] = rows.ravel()
    return cams, off


def _tracks(rng, n_pts, n_cams, lengths):
    """Distinct cameras per point, uniform without replacement, sorted."""
    cams = np.empty(int(lengths.sum()), dtype=np.int32)
    off = np.zeros(n_pts + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    for t in np.unique(lengths):
        idx = np.nonzero(lengths == t)[0]
        if 2 * t > n_cams:
            # long tracks in few cameras: rejection would rarely succeed, so
            # take the first t of a random permutation per row
            rows = np.argsort(rng.random((idx.size, n_cams)), axis=1)[:, :t].astype(np.int64)
            rows.sort(axis=1)
            pos = off[idx][:, None] + np.arange(t)[None, :]
            cams[pos.ravel()] = rows.ravel()
            continue
        rows = rng.integers(0, n_cams, size=(idx.size, t))
        rows.sort(axis=1)
        for _ in range(2000):
            bad = np.nonzero((np.diff(rows, axis=1) == 0).any(axis=1))[0] if t > 1 else np.zeros(0, int)
            if bad.size == 0:
                break
            redo = rng.integers(0, n_cams, size=(bad.size, t))
            redo.sort(axis=1)
            rows[bad] = redo
        else:
            # rows that drew a repeated camera every time: permutation draw
            bad = np.nonzero((np.diff(rows, axis=1) == 0).any(axis=1))[0]
            if bad.size:
                rows[bad] = np.sort(np.argsort(rng.random((bad.size, n_cams)), axis=1)[:, :t], axis=1)
        pos = off[idx][:, None] + np.arange(t)[None, :]
        cams[pos.ravel()] = rows.ravel()
    return cams, off


def generate(name=None, *, seed=1, n_cams=50, n_pts=5000, track="geometric", track_mean=9.0,
             outlier_frac=0.02, noise_px=0.5, outlier_px=50.0, f0=500.0, f_spread=0.05,
             k1_std=1e-3, k2_std=1e-5, rot_noise=0.01, t_noise=0.05, x_noise=0.05,
             radius=10.0, ball=3.0, cam_window=None, weights=None, img_size=None):
    """Return a scene dict (numpy arrays) in the reference's data model plus the
    ground truth.  ``name`` selects one of CONFIGS.

    weights="random": per-observation 2x2 information matrices (wx, wy, wxy)
    with wx, wy in [0.5, 2] and the correlation wxy / sqrt(wx wy) in
    [-0.6, 0.6]; the pixel noise is drawn with covariance noise_px^2 W^-1.
    img_size=(w, h): image sizes per camera (w + k mod 7, h), which set ihw2
    and the residual clamp 5000 / sqrt(ihw2) (rpbacore.cpp:87-90,
    rbacore.cpp:48)."""
    if name is not None:
        cfg = dict(CONFIGS[name])
        seed = cfg.pop("seed")
        return generate(None, seed=seed, **{**dict(track="geometric"), **cfg})
    rng = np.random.default_rng(seed)
    phi = 2.0 * np.pi * np.arange(n_cams) / n_cams
    centers = np.stack([radius * np.cos(phi), radius * np.sin(phi), np.zeros(n_cams)], axis=1)
    R_true = _look_at_origin(centers)
    t_true = -np.einsum("nij,nj->ni", R_true, centers)
    f_true = f0 * rng.uniform(1.0 - f_spread, 1.0 + f_spread, n_cams)
    k1 = rng.normal(0.0, k1_std, n_cams)
    k2 = rng.normal(0.0, k2_std, n_cams)

    # points uniform in the ball
    d = rng.normal(size=(n_pts, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    X_true = d * (ball * rng.uniform(0.0, 1.0, n_pts) ** (1.0 / 3.0))[:, None]

    if track == "uniform2_8":
        lengths = rng.integers(2, 9, n_pts)
    else:
        p = 1.0 / (track_mean - 1.0)
        lengths = 2 + rng.geometric(p, n_pts) - 1
    lengths = np.minimum(lengths, n_cams).astype(np.int64)
    if cam_window:
        lengths = np.minimum(lengths, cam_window)
        obs_cam, off = _window_tracks(rng, n_pts, n_cams, lengths, int(cam_window))
    else:
        obs_cam, off = _tracks(rng, n_pts, n_cams, lengths)
    obs_pt = np.repeat(np.arange(n_pts, dtype=np.int32), lengths)
    n_obs = obs_cam.size

    # observations from the true model
    P = np.einsum("nij,nj->ni", R_true[obs_cam], X_true[obs_pt]) + t_true[obs_cam]
    u = P[:, 0] / P[:, 2]
    v = P[:, 1] / P[:, 2]
    rho = u * u + v * v
    s = 1.0 + k1[obs_cam] * rho + k2[obs_cam] * rho * rho
    xg = np.stack([f_true[obs_cam] * s * u, f_true[obs_cam] * s * v], axis=1)
    if weights == "random":
        wa = rng.uniform(0.5, 2.0, n_obs)
        wb = rng.uniform(0.5, 2.0, n_obs)
        wc = rng.uniform(-0.6, 0.6, n_obs) * np.sqrt(wa * wb)
        obs_w = np.stack([wa, wb, wc], axis=1).astype(np.float32)
        # noise with covariance W^-1: x = L^-T z for W = L L^T
        l11 = np.sqrt(wa)
        l21 = wc / l11
        l22 = np.sqrt(wb - l21 * l21)
        z = rng.normal(0.0, noise_px, (n_obs, 2))
        e2 = z[:, 1] / l22
        e1 = (z[:, 0] - l21 * e2) / l11
        xg += np.stack([e1, e2], axis=1)
    else:
        obs_w = None
        xg += rng.normal(0.0, noise_px, xg.shape)
    n_out = int(round(outlier_frac * n_obs))
    out_idx = rng.choice(n_obs, n_out, replace=False)
    xg[out_idx] += rng.uniform(-outlier_px, outlier_px, (n_out, 2))
    is_outlier = np.zeros(n_obs, dtype=bool)
    is_outlier[out_idx] = True

    # initial values
    dR = rodrigues(rng.normal(0.0, rot_noise, (n_cams, 3)))
    R0 = np.einsum("nij,njk->nik", dR, R_true)
    aa0 = rot_to_aa(R0)
    t0 = (t_true + rng.normal(0.0, t_noise, (n_cams, 3))).astype(np.float32)
    X0 = (X_true + rng.normal(0.0, x_noise, (n_pts, 3))).astype(np.float32)
    R0f = rodrigues(aa0).astype(np.float32)
    cam_P = np.concatenate([R0f, t0[:, :, None]], axis=2)

    obs_xy = xg.astype(np.float32)  # reference xg (already -file xy)
    extra = {}
    if img_size is not None:
        extra["img_w"] = (int(img_size[0]) + np.arange(n_cams) % 7).astype(np.int32)
        extra["img_h"] = np.full(n_cams, int(img_size[1]), np.int32)
    return {**extra,
        "n_cams": n_cams, "n_pts": n_pts, "n_obs": n_obs,
        "obs_cam": obs_cam.astype(np.int32), "obs_pt": obs_pt,
        "obs_xy": obs_xy, "obs_w": obs_w,
        "cam_aa": aa0, "cam_t": t0,
        "cam_P": cam_P.astype(np.float32),
        "cam_f": f_true.astype(np.float32),
        "cam_d": np.stack([k1, k2], axis=1).astype(np.float32),
        "pts": X0,
        "truth": {"R": R_true, "t": t_true, "f": f_true, "k1": k1, "k2": k2, "X": X_true,
                  "outlier": is_outlier},
        "seed": seed,
    }


def rms_px(scene, out):
    """Per-coordinate RMS reprojection error in pixels, sqrt(sum |r|^2 / 2n),
    over the kept observations of a solved block (its output R|t, points and
    intrinsics), with the reference's camera model (rbacore.cpp:1238-1289;
    the residual is xg - p with p = f s (P X)_xy / (P X)_z)."""
    kept = np.asarray(out["obs_kept"]).astype(bool)
    cam = scene["obs_cam"][kept]
    P = np.asarray(out["cam_P"], dtype=np.float64).reshape(-1, 3, 4)[cam]
    X = np.asarray(out["pts"], dtype=np.float64).reshape(-1, 3)[scene["obs_pt"][kept]]
    e = np.einsum("nij,nj->ni", P[:, :, :3], X) + P[:, :, 3]
    z = e[:, 2]
    iz = np.where(np.abs(z) < 1e-3, np.where(z < 0, -1e3, 1e3), 1.0 / np.where(z == 0, 1.0, z))
    u, v = e[:, 0] * iz, e[:, 1] * iz
    rho = u * u + v * v
    d = np.asarray(out["cam_d"], dtype=np.float64).reshape(-1, 2)[cam]
    f = np.asarray(out["cam_f"], dtype=np.float64)[cam]
    sfac = 1.0 + d[:, 0] * rho + d[:, 1] * rho * rho
    xy = np.asarray(scene["obs_xy"], dtype=np.float64).reshape(-1, 2)[kept]
    r = xy - np.stack([f * sfac * u, f * sfac * v], axis=1)
    return float(np.sqrt((r * r).sum() / max(2 * r.shape[0], 1)))


def write_bal(path, scene):
    """BAL text in the layout InputBlock reads (rpba.cpp:195-279): header,
    observations `cam pt x y` (file x = -xg), cameras `a b c t0 t1 t2 f k1 k2`
    (k24out = 1 convention), points.  Floats are written with 9 significant
    digits and the angle-axis with 17, so reading back gives the same values."""
    with open(path, "w") as fh:
        fh.write(f"{scene['n_cams']} {scene['n_pts']} {scene['n_obs']}\n")
        xy = -scene["obs_xy"].astype(np.float32)
        lines = [f"{c} {p} {x:.9g} {y:.9g}\n" for c, p, x, y in
                 zip(scene["obs_cam"].tolist(), scene["obs_pt"].tolist(), xy[:, 0].tolist(), xy[:, 1].tolist())]
        fh.writelines(lines)
        for k in range(scene["n_cams"]):
            a = scene["cam_aa"][k]
            t = scene["cam_t"][k]
            fh.write(f"{a[0]:.17g} {a[1]:.17g} {a[2]:.17g} {t[0]:.9g} {t[1]:.9g} {t[2]:.9g} "
                     f"{scene['cam_f'][k]:.9g} {scene['cam_d'][k][0]:.9g} {scene['cam_d'][k][1]:.9g}\n")
        fh.writelines(f"{x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in scene["pts"].tolist())


def subscene(scene, n_cams, n_pts, seed=0):
    """Small scene cut from a bigger one (first n_cams cameras, points whose
    track keeps >= 2 of them), for quick tests."""
    keep_cam = scene["obs_cam"] < n_cams
    pts_idx = np.arange(n_pts)
    m = keep_cam & (scene["obs_pt"] < n_pts)
    cnt = np.bincount(scene["obs_pt"][m], minlength=n_pts)
    good = pts_idx[cnt >= 2]
    remap = -np.ones(scene["n_pts"], dtype=np.int64)
    remap[good] = np.arange(good.size)
    m &= remap[scene["obs_pt"]] >= 0
    out = {
        "n_cams": n_cams, "n_pts": int(good.size), "n_obs": int(m.sum()),
        "obs_cam": scene["obs_cam"][m].astype(np.int32),
        "obs_pt": remap[scene["obs_pt"][m]].astype(np.int32),
        "obs_xy": scene["obs_xy"][m], "obs_w": None,
        "cam_aa": scene["cam_aa"][:n_cams], "cam_t": scene["cam_t"][:n_cams],
        "cam_P": scene["cam_P"][:n_cams], "cam_f": scene["cam_f"][:n_cams],
        "cam_d": scene["cam_d"][:n_cams], "pts": scene["pts"][good],
        "seed": scene.get("seed", seed),
    }
    return out
