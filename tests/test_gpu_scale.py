"""The 13,682-camera scene (4.46M points, 29M observations, about 71M upper
blocks of S): the S-formed solve against the matrix-free one.

The two forms compute the same S d in different orders (S blocks from the
covisibility pairs, or W and Y records without forming S), so their first LM
iterations agree to rounding.  The oracle cannot run at this size (its S alone
is 46 GB); this comparison is what pins the sizes the small parity scenes never
reach: more than 2^31 elements of S, record arrays above 4 GiB, and launches
whose one-workgroup-per-block grid would exceed the 32-bit work-item count of
a dispatch (the off-diagonal blocks past that count were once left unformed).
Tolerances: the rule of tolerance.check_trajectory over two LM iterations."""
import sys

import numpy as np
import pytest

from rpba_amd import api, capi, scenes
from tolerance import check_trajectory

pytestmark = pytest.mark.gpu


def test_13682_cameras_formed_matches_matrix_free():
    sc = scenes.generate("c5_13682")
    traces = {}
    for schur in (0, 1):
        print(f"[c5_13682] schur={schur}: 2 LM iterations, n_obs={sc['n_obs']}", flush=True)
        sys.stdout.flush()
        ctx = api.Context(sc, capi.make_flags(robust=1, k2=1, k4=1, f=1, threads=1, schur=schur))
        rc, done, fin = ctx.run(2)
        assert rc == 0 and done == 2
        if schur == 0:
            sz = ctx.sizes()
            assert sz["blocks"] * 64 > 2**32, sz  # the size this test is for
        rc, res, tr = ctx.download()
        ctx.close()
        traces[schur] = tr
    a, b = traces[0], traces[1]
    check_trajectory(a["trace_sig0"], b["trace_sig0"], a["trace_pcg"], b["trace_pcg"],
                     a["trace_accept"], b["trace_accept"], a["trace_nrobs"], b["trace_nrobs"])
    assert np.all(np.isfinite(a["trace_sig0"]))
