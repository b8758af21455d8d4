"""ctypes mirror of include/rpba.h (the C ABI) and helpers that build a
``rpba_block`` from numpy arrays.

The structures follow the reference's ``CalcPLSA`` arguments
(rpba.h:693-703): observations in file order with their image and point
index (the tracks ``pointXXvio`` InputBlock builds, rpba.cpp:219-238), the
reference's ``xg`` = (-x_file, -y_file), the weights ``wx, wy, wxy``, the float
``PMat`` [R|t] per image, ``Camera.K(0,0)`` and ``distpara``, and ``XX``.
"""
import ctypes as C

import numpy as np

RPBA_OK = 0
RPBA_E_NO_OBS_IN_IMAGE = -1
RPBA_E_LOW_REDUNDANCY = -2
RPBA_E_NAN_RESULT = -3
RPBA_E_INVALID = -10
RPBA_E_UNSUPPORTED = -11
RPBA_E_SCRATCH = -12
RPBA_E_BUDGET = -13
RPBA_E_HIP = -20
RPBA_E_NOMEM = -21

_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class RpbaBlock(C.Structure):
    _fields_ = [
        ("n_cams", C.c_int32), ("n_pts", C.c_int32), ("n_obs", C.c_int32),
        ("obs_cam", _i32p), ("obs_pt", _i32p), ("obs_xy", _f32p), ("obs_w", _f32p),
        ("cam_P", _f32p), ("cam_f", _f32p), ("cam_d", _f32p), ("pts", _f32p),
        ("img_w", _i32p), ("img_h", _i32p),
        ("obs_kept", _u8p), ("pt_kept", _u8p),
    ]


class RpbaFlags(C.Structure):
    _fields_ = [
        ("robust", C.c_int32), ("exteval", C.c_int32), ("threads", C.c_int32),
        ("est_k2", C.c_int32), ("est_k4", C.c_int32), ("est_f", C.c_int32),
        ("k24out", C.c_int32), ("precision", C.c_int32), ("preconditioner", C.c_int32),
        ("force_global", C.c_int32), ("schur", C.c_int32), ("solver", C.c_int32), ("pc_omega", C.c_float),
    ]


class RpbaStats(C.Structure):
    _fields_ = [
        ("sigma_in", C.c_double * 2), ("sigma_out", C.c_double * 2),
        ("redundancy", C.c_int32 * 2), ("nobs", C.c_int32 * 2), ("nunknowns", C.c_int32 * 2),
        ("passes_done", C.c_int32), ("itt", C.c_int32), ("pcg_iters", C.c_int64),
        ("sig0_final", C.c_double), ("rms_px", C.c_double),
        ("n_obs_kept", C.c_int32), ("n_pts_kept", C.c_int32),
        ("exteval_sigma", C.c_double), ("exteval_nobs", C.c_int64), ("exteval_dof", C.c_int64),
        ("trace_cap", C.c_int32), ("trace_len", C.c_int32),
        ("trace_sig0", _f64p), ("trace_accept", _i32p), ("trace_pcg", _i32p), ("trace_nrobs", _i32p),
    ]


LOG_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_void_p)
LP_FEATURES = 48


class RpbaLpGraphs(C.Structure):
    _fields_ = [
        ("n_graphs", C.c_int32), ("n_nodes", C.c_int32), ("n_edges", C.c_int64),
        ("goff", _i32p), ("feat", _f32p), ("adj_off", _i32p), ("adj_nbr", _i32p), ("adj_w", _f32p),
        ("label", _f32p),
    ]
BATCH_LOG_FN = C.CFUNCTYPE(None, C.c_int32, C.c_char_p, C.c_void_p)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class Block:
    """Owns the numpy arrays behind one ``rpba_block`` (copies of the scene, so
    a solve can overwrite them in place the way CalcPLSA does)."""

    def __init__(self, scene):
        self.obs_cam = np.ascontiguousarray(scene["obs_cam"], dtype=np.int32)
        self.obs_pt = np.ascontiguousarray(scene["obs_pt"], dtype=np.int32)
        self.obs_xy = np.ascontiguousarray(scene["obs_xy"], dtype=np.float32).reshape(-1, 2)
        w = scene.get("obs_w")
        self.obs_w = None if w is None else np.ascontiguousarray(w, dtype=np.float32).reshape(-1, 3)
        self.cam_P = np.array(scene["cam_P"], dtype=np.float32).reshape(-1, 3, 4).copy()
        self.cam_f = np.array(scene["cam_f"], dtype=np.float32).copy()
        self.cam_d = np.array(scene["cam_d"], dtype=np.float32).reshape(-1, 2).copy()
        self.pts = np.array(scene["pts"], dtype=np.float32).reshape(-1, 3).copy()
        iw, ih = scene.get("img_w"), scene.get("img_h")
        self.img_w = None if iw is None else np.ascontiguousarray(iw, dtype=np.int32)
        self.img_h = None if ih is None else np.ascontiguousarray(ih, dtype=np.int32)
        self.n_cams = self.cam_f.shape[0]
        self.n_pts = self.pts.shape[0]
        self.n_obs = self.obs_cam.shape[0]
        self.obs_kept = np.zeros(self.n_obs, dtype=np.uint8)
        self.pt_kept = np.zeros(self.n_pts, dtype=np.uint8)
        self.c = RpbaBlock(
            self.n_cams, self.n_pts, self.n_obs,
            _ptr(self.obs_cam, C.c_int32), _ptr(self.obs_pt, C.c_int32),
            _ptr(self.obs_xy, C.c_float), _ptr(self.obs_w, C.c_float),
            _ptr(self.cam_P, C.c_float), _ptr(self.cam_f, C.c_float),
            _ptr(self.cam_d, C.c_float), _ptr(self.pts, C.c_float),
            _ptr(self.img_w, C.c_int32), _ptr(self.img_h, C.c_int32),
            _ptr(self.obs_kept, C.c_uint8), _ptr(self.pt_kept, C.c_uint8),
        )

    def result(self):
        return {
            "cam_P": self.cam_P.copy(), "cam_f": self.cam_f.copy(), "cam_d": self.cam_d.copy(),
            "pts": self.pts.copy(), "obs_kept": self.obs_kept.copy(), "pt_kept": self.pt_kept.copy(),
        }


def make_flags(robust=1, exteval=0, threads=1, k2=1, k4=1, f=1, k24out=1, precision=0,
               preconditioner=0, force_global=0, schur=0, solver=0, pc_omega=0.0):
    return RpbaFlags(robust, exteval, threads, k2, k4, f, k24out, precision, preconditioner, force_global, schur,
                     solver, pc_omega)


class Trace:
    """Caller-owned per-iteration trace arrays for ``rpba_stats``."""

    def __init__(self, cap=512):
        self.sig0 = np.zeros(cap, dtype=np.float64)
        self.accept = np.zeros(cap, dtype=np.int32)
        self.pcg = np.zeros(cap, dtype=np.int32)
        self.nrobs = np.zeros(cap, dtype=np.int32)
        self.stats = RpbaStats()
        self.stats.trace_cap = cap
        self.stats.trace_sig0 = _ptr(self.sig0, C.c_double)
        self.stats.trace_accept = _ptr(self.accept, C.c_int32)
        self.stats.trace_pcg = _ptr(self.pcg, C.c_int32)
        self.stats.trace_nrobs = _ptr(self.nrobs, C.c_int32)

    def as_dict(self):
        s = self.stats
        n = s.trace_len
        return {
            "sigma_in": list(s.sigma_in), "sigma_out": list(s.sigma_out),
            "redundancy": list(s.redundancy), "nobs": list(s.nobs), "nunknowns": list(s.nunknowns),
            "passes_done": s.passes_done, "itt": s.itt, "pcg_iters": s.pcg_iters,
            "sig0_final": s.sig0_final, "rms_px": s.rms_px,
            "n_obs_kept": s.n_obs_kept, "n_pts_kept": s.n_pts_kept,
            "exteval_sigma": s.exteval_sigma, "exteval_nobs": s.exteval_nobs, "exteval_dof": s.exteval_dof,
            "trace_sig0": self.sig0[:n].copy(), "trace_accept": self.accept[:n].copy(),
            "trace_pcg": self.pcg[:n].copy(), "trace_nrobs": self.nrobs[:n].copy(),
        }


class LogCollector:
    def __init__(self):
        self.lines = []
        self.fn = LOG_FN(self._cb)

    def _cb(self, line, user):
        self.lines.append(line.decode())
