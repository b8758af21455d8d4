"""Why the FP32 forms' first pass-1 PCG can miss the step-norm rule: the
4,585-camera scene solved by the oracle through pass 0 and the first LM
iteration of pass 1 (lambda = 1e-6, rbacore.cpp:365-376), in three
arithmetic modes of the linearisation (oracle_set_diag):

  0  plain FP64;
  1  the GPU's FP32 records emulated: Jacobians, Bc^T W, W = Bc^T W Bp and
     Y = W V^-1 rounded to float one by one, S summed in double from them;
  2  consistently rounded factors: the whitened Jacobians L^T Bc, L^T Bp
     (W = L L^T) rounded to float once, everything else formed from them in
     double.

The step-norm ratio m / m0 of every iteration of that first pass-1 PCG is
recorded (the PCG capped at 400 iterations), with the sig0 trajectory.

    python tests/golden/fp32_diagnostic.py MODE      # writes fp32_diag_<MODE>.json
"""
import ctypes as C
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

CAP = 400


def run(mode, scene="c3_4585", pass0_iters=4):
    sc = scenes.generate(scene)
    buf = np.zeros(CAP, dtype=np.float64)
    # PCG calls are counted from here: pass 0 runs pass0_iters of them
    oracle.lib().oracle_set_diag(mode, CAP, pass0_iters + 1, pass0_iters,
                                 buf.ctypes.data_as(C.POINTER(C.c_double)), CAP)
    t = time.time()
    rc, out, st, log = oracle.solve(sc, capi.make_flags(robust=1), fsites=0)
    n = oracle.lib().oracle_diag_trace_len()
    oracle.lib().oracle_set_diag(0, 0, 0, -1, None, 0)
    res = {"scene": scene, "mode": mode, "status": rc, "trace_sig0": [float(x) for x in st["trace_sig0"]],
           "trace_pcg": [int(x) for x in st["trace_pcg"]], "pass1_first_pcg_m_over_m0": buf[:n].tolist(),
           "met_rule": bool(n and buf[n - 1] * 1e5 < 1.0), "min_ratio": float(buf[:n].min()) if n else None,
           "seconds": round(time.time() - t, 1)}
    with open(os.path.join(HERE, f"fp32_diag_{mode}.json"), "w") as fh:
        json.dump(res, fh)
    print(mode, "pcg", res["trace_pcg"], "sig0", res["trace_sig0"], "met", res["met_rule"], "min", res["min_ratio"],
          flush=True)


if __name__ == "__main__":
    run(int(sys.argv[1]))
