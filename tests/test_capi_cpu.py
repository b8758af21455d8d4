"""C ABI checks that need no GPU: the library loads and exports every entry
point include/rpba.h declares; BAL and `parameters` I/O (host only)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import rpba_amd
from rpba_amd import capi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "rpba.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(rpba_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = rpba_amd.lib()
    names = _header_functions()
    assert len(names) >= 14
    assert sorted(rpba_amd.EXPORTS) == names
    for n in names:
        assert hasattr(L, n), n
    assert L.rpba_version() == 1


def test_bal_round_trip(tmp_path):
    sc = scenes.generate(None, seed=21, n_cams=7, n_pts=90, track="uniform2_8", outlier_frac=0.0)
    p = tmp_path / "p.txt"
    scenes.write_bal(str(p), sc)
    L = rpba_amd.lib()
    blk = capi.RpbaBlock()
    assert L.rpba_read_bal(str(p).encode(), 1, C.byref(blk)) == 0
    try:
        assert (blk.n_cams, blk.n_pts, blk.n_obs) == (7, 90, sc["n_obs"])
        xy = np.ctypeslib.as_array(blk.obs_xy, shape=(blk.n_obs * 2,)).reshape(-1, 2)
        assert np.array_equal(xy, sc["obs_xy"])  # 9 digits: float values round-trip exactly
        cam = np.ctypeslib.as_array(blk.obs_cam, shape=(blk.n_obs,))
        assert np.array_equal(cam, sc["obs_cam"])
        P = np.ctypeslib.as_array(blk.cam_P, shape=(7 * 12,)).reshape(7, 3, 4)
        assert np.array_equal(P[:, :, 3], sc["cam_t"])
        assert np.abs(P[:, :, :3] - sc["cam_P"][:, :, :3]).max() <= 2 ** -23
        X = np.ctypeslib.as_array(blk.pts, shape=(90 * 3,)).reshape(-1, 3)
        assert np.array_equal(X, sc["pts"])
        # k24out = 0 multiplies the distortion by f^2, f^4 in float (rpba.cpp:259-261)
        blk0 = capi.RpbaBlock()
        assert L.rpba_read_bal(str(p).encode(), 0, C.byref(blk0)) == 0
        d0 = np.ctypeslib.as_array(blk0.cam_d, shape=(14,)).reshape(-1, 2)
        f = sc["cam_f"]
        assert np.array_equal(d0[:, 0], sc["cam_d"][:, 0] * f * f)
        # write then read back: same observations, 8 significant digits
        q = tmp_path / "q.txt"
        assert L.rpba_write_bal(str(q).encode(), C.byref(blk), 1) == 0
        blk2 = capi.RpbaBlock()
        assert L.rpba_read_bal(str(q).encode(), 1, C.byref(blk2)) == 0
        xy2 = np.ctypeslib.as_array(blk2.obs_xy, shape=(blk2.n_obs * 2,)).reshape(-1, 2)
        assert np.allclose(xy2, xy, rtol=1e-7, atol=1e-5)
        P2 = np.ctypeslib.as_array(blk2.cam_P, shape=(7 * 12,)).reshape(7, 3, 4)
        # acos(trace) loses digits near 180 degrees (rpba.cpp:432): ~1e-6 there
        assert np.allclose(P2, P, atol=5e-6)
        L.rpba_free_block(C.byref(blk0))
        L.rpba_free_block(C.byref(blk2))
    finally:
        L.rpba_free_block(C.byref(blk))


def test_parameters_reader(tmp_path):
    L = rpba_amd.lib()
    p = tmp_path / "parameters"
    p.write_text("Robust = 1\nExteval = 0\nThreads = 0\n\n\nk2 = 1\nk4 = 0\nf = 1\n\nk24out = 0\n\n\n"
                 "InputDir = ./961\nFile = problem-961-187103-pre.txt\n")
    fl = capi.RpbaFlags()
    d = C.create_string_buffer(256)
    f = C.create_string_buffer(256)
    assert L.rpba_read_parameters(str(p).encode(), C.byref(fl), d, 256, f, 256) == 0
    assert (fl.robust, fl.exteval, fl.threads, fl.est_k2, fl.est_k4, fl.est_f, fl.k24out) == (1, 0, 0, 1, 0, 1, 0)
    assert d.value == b"./961" and f.value == b"problem-961-187103-pre.txt"
    p.write_text("Robust = 1\n")  # every key is required (rpba.cpp:56-64)
    assert L.rpba_read_parameters(str(p).encode(), C.byref(fl), d, 256, f, 256) == capi.RPBA_E_INVALID


def test_invalid_arguments_are_rejected_without_gpu():
    L = rpba_amd.lib()
    assert L.rpba_calc_plsa(None, None, capi.LOG_FN(0), None, None) == capi.RPBA_E_INVALID


def test_unknown_form_selectors_are_rejected_without_gpu(tiny_scene):
    """precision outside {0 .. 3} and schur outside {0, 1, 2} fail before any
    device call, and so do the whitened-Jacobian forms (precision 2, 3) with
    S formed, the dense solver or a corrected preconditioner: their records
    are read by the matrix-free product only."""
    from rpba_amd import api
    for kw in ({"precision": 4}, {"precision": -1}, {"schur": 3}, {"schur": -1},
               {"precision": 2, "schur": 0}, {"precision": 3, "schur": 0}, {"precision": 2, "schur": 2, "solver": 1},
               {"precision": 2, "schur": 2, "preconditioner": 1, "pc_omega": 0.5},
               {"precision": 3, "schur": 1, "preconditioner": 2}):
        rc, _, _, _ = api.calc_plsa(tiny_scene, capi.make_flags(**kw))
        assert rc == capi.RPBA_E_INVALID, kw


def test_window_tracks_are_banded_and_distinct():
    """The camera-window generator (sparse, banded covisibility) gives every
    point distinct cameras inside one window of consecutive cameras."""
    sc = scenes.generate(None, seed=51, n_cams=60, n_pts=3000, track="uniform2_8", outlier_frac=0.02, cam_window=10)
    cam, pt = sc["obs_cam"], sc["obs_pt"]
    off = np.r_[0, np.cumsum(np.bincount(pt, minlength=sc["n_pts"]))]
    for i in range(sc["n_pts"]):
        c = np.sort(cam[off[i]:off[i + 1]])
        assert len(c) >= 2 and len(np.unique(c)) == len(c)
        # on the circle of 60 cameras, all of them fit in 10 consecutive ones
        gaps = np.diff(np.r_[c, c[0] + 60])
        assert 60 - gaps.max() <= 9


def test_batch_form_selectors_are_rejected_without_gpu(tiny_scene):
    """The batched blocks take the reference's PCG or the dense solve with S
    formed, FP64 or FP32 records and block-Jacobi only: any other camera solve,
    product, record form or preconditioner fails before any device call."""
    from rpba_amd import api
    L = rpba_amd.lib()
    for kw in ({"solver": 2}, {"solver": -1}, {"schur": 1}, {"schur": 2}, {"precision": 2}, {"precision": 3},
               {"precision": -1}, {"preconditioner": 1, "pc_omega": 0.5}, {"preconditioner": 2}):
        with pytest.raises(RuntimeError, match=f"status {capi.RPBA_E_INVALID}"):
            api.Batch([tiny_scene, tiny_scene], capi.make_flags(**kw))
    h = C.c_void_p()
    fl = capi.make_flags()
    assert L.rpba_batch_create(C.byref(h), 0, None, C.byref(fl), capi.BATCH_LOG_FN(0), None) == capi.RPBA_E_INVALID
