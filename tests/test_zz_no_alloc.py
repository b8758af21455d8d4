"""rpba_exteval_device (Exteval from device-resident tracks, no device
allocation) against the allocating entry point rpba_calc_plsa with
exteval = 1, on two calls in a row, the second on a smaller input.  Both run
the same kernels, so the results are equal to the last bit."""
import ctypes as C

import numpy as np
import pytest

import rpba_amd
from rpba_amd import api, capi, scenes

pytestmark = pytest.mark.gpu


def _tracks(sc):
    """Point-major CSR of the observations (file order within a track), as
    InputBlock's pointXXvio lists them (rpba.cpp:219-238)."""
    order = np.argsort(sc["obs_pt"], kind="stable")
    cnt = np.bincount(sc["obs_pt"], minlength=sc["n_pts"])
    toff = np.zeros(sc["n_pts"] + 1, dtype=np.int32)
    np.cumsum(cnt, out=toff[1:])
    w = None if sc["obs_w"] is None else np.ascontiguousarray(sc["obs_w"][order], dtype=np.float32)
    return toff, sc["obs_cam"][order].astype(np.int32), np.ascontiguousarray(sc["obs_xy"][order]), w


class _Hip:
    """The caller's own device memory, through the HIP runtime the library
    itself links (torch ships another HIP runtime, which must not share the
    process's device with it)."""

    def __init__(self):
        rpba_amd.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _device_exteval(hip, sc, flags, img=None):
    toff, tcam, txy, tw = _tracks(sc)
    out0 = np.full(3, -1.0)
    d = [hip.put(x) for x in (toff, tcam, txy, tw, sc["cam_P"].reshape(-1), sc["cam_f"],
                              sc["cam_d"].reshape(-1), sc["pts"].reshape(-1),
                              None if img is None else img[0], None if img is None else img[1], out0)]
    rc = rpba_amd.lib().rpba_exteval_device(sc["n_cams"], sc["n_pts"], *d[:10], flags.est_k2, flags.est_k4,
                                            flags.est_f, d[10], None)
    assert hip.L.hipDeviceSynchronize() == 0
    out = hip.get(d[10], out0)
    hip.free()
    return rc, out


def _with_weights(sc, seed):
    rng = np.random.default_rng(seed)
    n = sc["n_obs"]
    a = rng.uniform(0.5, 2.0, n)
    b = rng.uniform(0.5, 2.0, n)
    c = rng.uniform(-0.4, 0.4, n) * np.sqrt(a * b)
    sc = dict(sc)
    sc["obs_w"] = np.stack([a, b, c], axis=1).astype(np.float32)
    return sc


def test_no_alloc_exteval_matches_allocating_entry_point():
    big = _with_weights(scenes.generate(None, seed=61, n_cams=40, n_pts=9000, track="uniform2_8",
                                        outlier_frac=0.02), 1)
    small = scenes.generate(None, seed=62, n_cams=12, n_pts=700, track="uniform2_8", outlier_frac=0.02)
    img = (np.full(40, 640, np.int32), np.full(40, 480, np.int32))
    hip = _Hip()
    for sc, im, name in ((big, img, "40 cameras, weights, image sizes"), (small, None, "12 cameras")):
        print(f"[no-alloc exteval] {name}: n_pts={sc['n_pts']} n_obs={sc['n_obs']}", flush=True)
        fl = capi.make_flags(exteval=1, k4=0)
        scr = dict(sc)
        if im is not None:
            scr["img_w"], scr["img_h"] = im
        rc0, _, st, log = api.calc_plsa(scr, fl)
        assert rc0 == 0 and log[0].startswith("ITTTT: ")
        rc, out = _device_exteval(hip, sc, fl, im)
        assert rc == 0
        assert out[1] == st["exteval_nobs"] and out[2] == st["exteval_dof"]
        assert out[0] == st["exteval_sigma"], (out[0], st["exteval_sigma"])


def test_no_alloc_exteval_rejects_oversize_input():
    n = rpba_amd.lib().rpba_exteval_device_max_points()
    assert n == 8388608
    z = C.c_void_p(16)  # never dereferenced: the size check comes first
    rc = rpba_amd.lib().rpba_exteval_device(3, n + 1, z, z, z, None, z, z, z, z, None, None, 1, 1, 1, z, None)
    assert rc == capi.RPBA_E_SCRATCH
