"""The `rpba` host driver (rpba_amd/bin/rpba): the reference's main
(rpba.cpp:45-176) on the device solve.  It reads ./parameters and the BAL
file, prints the reference's lines, and with the optional keys writes the
solved block (`Output = 1`) and the point cloud (`Cloud = 1`)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import rpba_amd
from rpba_amd import api, capi, scenes

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(rpba_amd.LIB_PATH), "bin", "rpba")


def _params(tmp, exteval=0, extra=""):
    (tmp / "parameters").write_text(
        f"Robust = 1\nExteval = {exteval}\nThreads = 1\nk2 = 1\nk4 = 1\nf = 1\nk24out = 1\n"
        f"InputDir = {tmp}\nFile = scene.txt\n{extra}")


def _run(tmp):
    print(f"[driver] {BIN} in {tmp}", flush=True)
    sys.stdout.flush()
    r = subprocess.run([BIN], cwd=tmp, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout


def test_driver_solve_output_and_cloud(tmp_path):
    sc = scenes.generate(None, seed=44, n_cams=12, n_pts=500, track="uniform2_8", outlier_frac=0.03)
    scenes.write_bal(str(tmp_path / "scene.txt"), sc)
    _params(tmp_path, extra="Output = 1\nCloud = 1\n")
    rc, out = _run(tmp_path)
    assert rc == 0, out
    lines = out.splitlines()
    assert lines[0] == f"Inputfile: {tmp_path}/scene.txt"
    assert f"Cameras: 12  Points: 500 Observations {sc['n_obs']}" in out
    assert sum(1 for l in lines if l.startswith("Sigma0:")) == 2
    assert "Bundle adjustment runtime:" in out
    # the written block equals the library's solve of the same file
    L = rpba_amd.lib()
    blk = capi.RpbaBlock()
    assert L.rpba_read_bal(str(tmp_path / "scene.txt").encode(), 1, C.byref(blk)) == 0
    res = capi.RpbaBlock()
    assert L.rpba_read_bal(str(tmp_path / "out-scene.txt").encode(), 1, C.byref(res)) == 0
    try:
        sc2 = {"n_cams": blk.n_cams, "n_pts": blk.n_pts, "n_obs": blk.n_obs,
               "obs_cam": np.ctypeslib.as_array(blk.obs_cam, shape=(blk.n_obs,)).copy(),
               "obs_pt": np.ctypeslib.as_array(blk.obs_pt, shape=(blk.n_obs,)).copy(),
               "obs_xy": np.ctypeslib.as_array(blk.obs_xy, shape=(2 * blk.n_obs,)).reshape(-1, 2).copy(),
               "cam_P": np.ctypeslib.as_array(blk.cam_P, shape=(12 * blk.n_cams,)).reshape(-1, 3, 4).copy(),
               "cam_f": np.ctypeslib.as_array(blk.cam_f, shape=(blk.n_cams,)).copy(),
               "cam_d": np.ctypeslib.as_array(blk.cam_d, shape=(2 * blk.n_cams,)).reshape(-1, 2).copy(),
               "pts": np.ctypeslib.as_array(blk.pts, shape=(3 * blk.n_pts,)).reshape(-1, 3).copy()}
        rc2, o2, st2, _ = api.calc_plsa(sc2, capi.make_flags(robust=1, threads=1))
        assert rc2 == 0
        kept = o2["obs_kept"].astype(bool)
        assert res.n_obs == int(kept.sum()) and res.n_cams == 12
        pts_out = np.ctypeslib.as_array(res.pts, shape=(3 * res.n_pts,)).reshape(-1, 3)
        # OutputBlock writes %.8g; the cloud holds the kept points
        cloud = np.loadtxt(tmp_path / "cloud.txt").reshape(-1, 3)
        pk = o2["pt_kept"].astype(bool)
        assert cloud.shape[0] == int(pk.sum())
        assert np.allclose(cloud, o2["pts"][pk], rtol=1e-7, atol=1e-6)
        assert f"Counter: {int((~pk).sum())}" in out
        assert pts_out.shape[0] == res.n_pts
    finally:
        L.rpba_free_block(C.byref(blk))
        L.rpba_free_block(C.byref(res))


def test_driver_exteval(tmp_path):
    sc = scenes.generate(None, seed=45, n_cams=8, n_pts=300, track="uniform2_8", outlier_frac=0.0)
    scenes.write_bal(str(tmp_path / "scene.txt"), sc)
    _params(tmp_path, exteval=1)
    rc, out = _run(tmp_path)
    assert rc == 0, out
    itt = [l for l in out.splitlines() if l.startswith("ITTTT: ")]
    assert len(itt) == 1 and float(itt[0].split()[1]) > 0
    assert not (tmp_path / "out-scene.txt").exists()


@pytest.mark.parametrize("extra", ["Solver = 1\n", "Preconditioner = 1\nPcOmega = 0.7\n", "Preconditioner = 2\n",
                                   "Precision = 1\nSchur = 1\n"])
def test_driver_form_keys(tmp_path, extra):
    """The forms' keys (ignored by the reference) select the same form as the
    library flags: the driver's printed Sigma0 lines equal the library's."""
    sc = scenes.generate(None, seed=46, n_cams=10, n_pts=400, track="uniform2_8", outlier_frac=0.02)
    scenes.write_bal(str(tmp_path / "scene.txt"), sc)
    _params(tmp_path, extra=extra + "Output = 1\n")
    rc, out = _run(tmp_path)
    assert rc == 0, out
    sig = [l for l in out.splitlines() if l.startswith("Sigma0:")]
    kv = dict(l.split("=") for l in extra.strip().splitlines())
    kv = {k.strip(): v.strip() for k, v in kv.items()}
    fl = capi.make_flags(robust=1, threads=1, precision=int(kv.get("Precision", 0)), schur=int(kv.get("Schur", 0)),
                         solver=int(kv.get("Solver", 0)), preconditioner=int(kv.get("Preconditioner", 0)),
                         pc_omega=float(kv.get("PcOmega", 0)))
    L = rpba_amd.lib()
    blk = capi.RpbaBlock()
    assert L.rpba_read_bal(str(tmp_path / "scene.txt").encode(), 1, C.byref(blk)) == 0
    try:
        sc2 = {"n_cams": blk.n_cams, "n_pts": blk.n_pts, "n_obs": blk.n_obs,
               "obs_cam": np.ctypeslib.as_array(blk.obs_cam, shape=(blk.n_obs,)).copy(),
               "obs_pt": np.ctypeslib.as_array(blk.obs_pt, shape=(blk.n_obs,)).copy(),
               "obs_xy": np.ctypeslib.as_array(blk.obs_xy, shape=(2 * blk.n_obs,)).reshape(-1, 2).copy(),
               "cam_P": np.ctypeslib.as_array(blk.cam_P, shape=(12 * blk.n_cams,)).reshape(-1, 3, 4).copy(),
               "cam_f": np.ctypeslib.as_array(blk.cam_f, shape=(blk.n_cams,)).copy(),
               "cam_d": np.ctypeslib.as_array(blk.cam_d, shape=(2 * blk.n_cams,)).reshape(-1, 2).copy(),
               "pts": np.ctypeslib.as_array(blk.pts, shape=(3 * blk.n_pts,)).reshape(-1, 3).copy()}
    finally:
        L.rpba_free_block(C.byref(blk))
    ctx = api.Context(sc2, fl)
    if fl.preconditioner == 2:
        from rpba_amd import lp_train
        ctx.set_lp(api.LpModel.load(lp_train.WEIGHTS))
    rc2, n, fin = ctx.run(-1)
    assert rc2 == 0
    lines = ctx.logc.lines
    ctx.close()
    assert [l.strip() for l in lines if l.startswith("Sigma0:")] == [l.strip() for l in sig]
