"""Writes c4_blocks.json: the oracle's solve (plain FP64, robust, exact camera
solve, solver = 1) of each of the 64 independent blocks of the batched-blocks
workload (scenes.BATCH_CONFIG: 20-400 cameras log-uniform, 8.5M observations
in all), one block at a time, as the reference solves its partition blocks one
LSA per thread (rpbacore.cpp:219-235); with the argument pcg, c4_blocks_pcg.json:
the same with the reference's camera solve, block-Jacobi SparsePCG
(rbacore.cpp:1807-1914, solver = 0).  The GPU's batched solve of all blocks at
once is held to these records (tests/test_gpu_c4_blocks.py).

    python tests/golden/make_golden_blocks.py [workers] [pcg]
"""
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_golden import digest  # noqa: E402
from oracle import oracle  # noqa: E402
from rpba_amd import capi, scenes  # noqa: E402

FLAGS = dict(robust=1, k2=1, k4=1, f=1, threads=1, solver=1)
if len(sys.argv) > 2 and sys.argv[2] == "pcg":
    FLAGS["solver"] = 0
OUT = "c4_blocks.json" if FLAGS["solver"] == 1 else "c4_blocks_pcg.json"


def one(b):
    sc = scenes.batch_block(b, **scenes.BATCH_CONFIG)
    t = time.time()
    rc, out, st, log = oracle.solve(sc, capi.make_flags(**FLAGS), fsites=0)
    rec = {
        "block": b, "n_cams": int(sc["n_cams"]), "n_pts": int(sc["n_pts"]), "n_obs": int(sc["n_obs"]),
        "status": rc, "itt": st["itt"], "passes_done": st["passes_done"],
        "trace_sig0": [float(x) for x in st["trace_sig0"]],
        "trace_accept": [int(x) for x in st["trace_accept"]],
        "trace_nrobs": [int(x) for x in st["trace_nrobs"]],
        "trace_pcg": [int(x) for x in st["trace_pcg"]],
        "sigma_in": list(st["sigma_in"]), "sigma_out": list(st["sigma_out"]),
        "redundancy": list(st["redundancy"]), "nobs": list(st["nobs"]), "nunknowns": list(st["nunknowns"]),
        "obs_kept_sha256": digest(out["obs_kept"]), "pt_kept_sha256": digest(out["pt_kept"]),
        "cam_f": [float(x) for x in out["cam_f"]],
        "pts_mean": [float(x) for x in out["pts"].astype(np.float64).mean(axis=0)],
        "log": log, "oracle_seconds": round(time.time() - t, 2),
    }
    print("block", b, "cameras", sc["n_cams"], "itt", st["itt"], "seconds", rec["oracle_seconds"], flush=True)
    return rec


if __name__ == "__main__":
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    nb = scenes.BATCH_CONFIG["n_blocks"]
    # largest blocks first, so the pool finishes together
    rng = np.random.default_rng(scenes.BATCH_CONFIG["seed"])
    c = scenes.BATCH_CONFIG
    sizes = np.exp(rng.uniform(np.log(c["cam_lo"]), np.log(c["cam_hi"]), nb))
    order = sorted(range(nb), key=lambda b: -sizes[b])
    with Pool(workers) as pool:
        recs = pool.map(one, order, chunksize=1)
    recs.sort(key=lambda r: r["block"])
    with open(os.path.join(HERE, OUT), "w") as fh:
        json.dump({"config": scenes.BATCH_CONFIG, "flags": FLAGS, "blocks": recs}, fh)
