"""Measures how far rounding alone moves the oracle's LM trajectory: the same
scene solved by the oracle in its reference summation order and in variants
that change only the rounding (the S d row sums in reverse column order;
the Schur pair products as fused multiply-adds; both).  The result is the
rounding spread two correct FP64 implementations of the reference can show,
the basis of the trajectory bar in tests/tolerance.py.

    python tests/golden/oracle_spread.py [case ...]     # writes spread_<case>.json

A case c4_block<b> is block b of the batched-blocks workload
(scenes.BATCH_CONFIG) with the reference's camera solve, SparsePCG.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import make_golden  # noqa: E402
from oracle import oracle  # noqa: E402
from rpba_amd import capi, scenes  # noqa: E402

VARIANTS = {1: "S d rows in reverse column order", 2: "fused Schur pair products", 3: "both"}


def case_spec(case):
    return make_golden.CASES[case] if case in make_golden.CASES else make_golden.LARGE[case]


def case_scene_flags(case):
    if case.startswith("c4_block"):
        sc = scenes.batch_block(int(case[len("c4_block"):]), **scenes.BATCH_CONFIG)
        return sc, capi.make_flags(robust=1, k2=1, k4=1, f=1, threads=1, solver=0)
    c = case_spec(case)
    return make_golden.make_scene(c["scene"]), capi.make_flags(**c["flags"])


def solve(sc, flags, order):
    with oracle.order(order):
        t = time.time()
        rc, out, st, log = oracle.solve(sc, flags, fsites=0)
        return rc, out, st, time.time() - t


def measure(case, variants=(1, 2, 3)):
    sc, flags = case_scene_flags(case)
    rc0, out0, st0, dt0 = solve(sc, flags, 0)
    res = {"case": case, "base": {"itt": st0["itt"], "trace_sig0": [float(x) for x in st0["trace_sig0"]],
                                  "trace_pcg": [int(x) for x in st0["trace_pcg"]],
                                  "sigma_out": list(st0["sigma_out"]), "seconds": round(dt0, 1)},
           "variants": {}}
    for v in variants:
        rc, out, st, dt = solve(sc, flags, v)
        a, b = np.asarray(st["trace_sig0"]), np.asarray(st0["trace_sig0"])
        n = min(a.size, b.size)
        rel = (np.abs(a[:n] - b[:n]) / np.abs(b[:n])).tolist()
        res["variants"][str(v)] = {
            "what": VARIANTS[v], "itt": st["itt"], "rel_sig0_per_iter": rel,
            "trace_pcg": [int(x) for x in st["trace_pcg"]],
            "same_decisions": bool(np.array_equal(st["trace_accept"], st0["trace_accept"])),
            "same_labels": bool(np.array_equal(out["obs_kept"], out0["obs_kept"])),
            "sigma_out_rel": [abs(x - y) / y for x, y in zip(st["sigma_out"], st0["sigma_out"])],
            "rms_rel": abs(scenes.rms_px(sc, out) - scenes.rms_px(sc, out0)) / scenes.rms_px(sc, out0),
            "cam_f_rel_max": float(np.max(np.abs(out["cam_f"].astype(np.float64) - out0["cam_f"])
                                          / np.abs(out0["cam_f"].astype(np.float64)))),
            "pts_diff_max_over_extent": float(np.abs(out["pts"].astype(np.float64) - out0["pts"]).max()
                                              / np.abs(out0["pts"].astype(np.float64)).max()),
            "seconds": round(dt, 1)}
        print(case, "variant", v, "max rel sig0 %.3g" % max(rel), "pcg", res["variants"][str(v)]["trace_pcg"],
              flush=True)
    with open(os.path.join(HERE, f"spread_{case}.json"), "w") as fh:
        json.dump(res, fh, indent=1)


if __name__ == "__main__":
    for case in (sys.argv[1:] or ["c1_plain", "c1_robust"]):
        measure(case)
