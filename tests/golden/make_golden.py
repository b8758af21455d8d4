"""Writes the golden trajectories in this folder from the CPU oracle (plain
FP64 mode) on seeded synthetic scenes.  The fixtures pin the GPU build: a GPU
solve of the same scene must reproduce the per-iteration sig0 (1e-6
relative), the accept decisions, nrobs and the robust labels exactly.

    python tests/golden/make_golden.py [name ...]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from rpba_amd import capi, scenes  # noqa: E402

CASES = {
    "c1_robust": dict(scene=("c1_50", {}), flags=dict(robust=1)),
    "c1_plain": dict(scene=("c1_50", {}), flags=dict(robust=0)),
    "c1_robust_k2f": dict(scene=("c1_50", {}), flags=dict(robust=1, k4=0)),
    "small_robust_nointr": dict(scene=(None, dict(seed=51, n_cams=20, n_pts=1500, track="uniform2_8",
                                                  outlier_frac=0.05)), flags=dict(robust=1, k2=0, k4=0, f=0)),
    "c2_robust": dict(scene=("c2_961", {}), flags=dict(robust=1)),
}

# Large scenes: the fixture also keeps a strided sample of the solved cameras
# and points, so the GPU forms' parameter errors can be stated against it.
LARGE = {
    "c3_robust": dict(scene=("c3_4585", {}), flags=dict(robust=1), cam_stride=23, pt_stride=997),
}


def make_scene(spec):
    name, kw = spec
    return scenes.generate(name) if name else scenes.generate(None, **kw)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(case):
    c = CASES[case] if case in CASES else LARGE[case]
    sc = make_scene(c["scene"])
    flags = capi.make_flags(**c["flags"])
    t = time.time()
    rc, out, st, log = oracle.solve(sc, flags, fsites=0)
    dt = time.time() - t
    removed = np.nonzero(out["obs_kept"] == 0)[0]
    fix = {
        "case": case, "scene": list(c["scene"][:1]) + [c["scene"][1]], "flags": c["flags"], "status": rc,
        "n_obs": int(sc["n_obs"]), "itt": st["itt"], "passes_done": st["passes_done"],
        "trace_sig0": [float(x) for x in st["trace_sig0"]],
        "trace_accept": [int(x) for x in st["trace_accept"]],
        "trace_nrobs": [int(x) for x in st["trace_nrobs"]],
        "trace_pcg": [int(x) for x in st["trace_pcg"]],
        "sigma_in": list(st["sigma_in"]), "sigma_out": list(st["sigma_out"]),
        "redundancy": list(st["redundancy"]), "nobs": list(st["nobs"]), "nunknowns": list(st["nunknowns"]),
        "n_obs_kept": int(out["obs_kept"].sum()), "n_pts_kept": int(out["pt_kept"].sum()),
        "obs_kept_sha256": digest(out["obs_kept"]), "pt_kept_sha256": digest(out["pt_kept"]),
        "removed_first": [int(x) for x in removed[:50]],
        "pts_mean": [float(x) for x in out["pts"].astype(np.float64).mean(axis=0)],
        "pts_abs_sum": float(np.abs(out["pts"].astype(np.float64)).sum()),
        "cam_f_first": [float(x) for x in out["cam_f"][:8]],
        "cam_t_first": [float(x) for x in out["cam_P"][:4, :, 3].ravel()],
        "log": log, "oracle_seconds": round(dt, 2),
    }
    if "cam_stride" in c:
        cs, ps = c["cam_stride"], c["pt_stride"]
        fix["cam_stride"], fix["pt_stride"] = cs, ps
        fix["cam_P_sample"] = out["cam_P"][::cs].astype(np.float64).ravel().tolist()
        fix["cam_f_sample"] = out["cam_f"][::cs].astype(np.float64).tolist()
        fix["cam_d_sample"] = out["cam_d"][::cs].astype(np.float64).ravel().tolist()
        fix["pts_sample"] = out["pts"][::ps].astype(np.float64).ravel().tolist()
        fix["sig0_final"] = float(st["sig0_final"])
        fix["rms_px"] = scenes.rms_px(sc, out)
    with open(os.path.join(HERE, f"{case}.json"), "w") as fh:
        json.dump(fix, fh, indent=1)
    print(case, "itt", st["itt"], "seconds", round(dt, 1), flush=True)


if __name__ == "__main__":
    for case in (sys.argv[1:] or [k for k in CASES if k != "c2_robust"]):
        run(case)
