"""Batched independent blocks (rpba_batch_*): every block's result against
the oracle's solve of that block alone with the same camera solve (solver 1,
exact; solver 0, the reference's block-Jacobi PCG segmented by block), under
the trajectory rule of the parity tests; hipGraph replay against eager
launches (bitwise); reset and re-run (bitwise)."""
import numpy as np
import pytest

from oracle import oracle
from rpba_amd import api, capi, scenes
from tolerance import check_trajectory

pytestmark = pytest.mark.gpu


def _blocks(seed, sizes, outliers=0.03):
    return [scenes.generate(None, seed=seed + i, n_cams=nc, n_pts=npt, track="uniform2_8", outlier_frac=outliers)
            for i, (nc, npt) in enumerate(sizes)]


def _compare_block(sc, flags, got, tag):
    rc1, out1, st1, log1 = got
    rc0, out0, st0, log0 = oracle.solve(sc, flags, fsites=0)
    print(f"[{tag}] n_cams={sc['n_cams']} itt oracle {st0['itt']} batch {st1['itt']}", flush=True)
    assert rc1 == rc0
    assert st1["passes_done"] == st0["passes_done"]
    assert st1["itt"] == st0["itt"], (st0["trace_sig0"], st1["trace_sig0"])
    check_trajectory(st1["trace_sig0"], st0["trace_sig0"], st1["trace_pcg"], st0["trace_pcg"],
                     st1["trace_accept"], st0["trace_accept"], st1["trace_nrobs"], st0["trace_nrobs"])
    for p in range(st0["passes_done"]):
        for key in ("sigma_in", "sigma_out"):
            assert abs(st1[key][p] - st0[key][p]) <= 1e-6 * abs(st0[key][p])
        for key in ("redundancy", "nobs", "nunknowns"):
            assert st1[key][p] == st0[key][p]
    assert np.array_equal(out1["obs_kept"], out0["obs_kept"])
    assert np.array_equal(out1["pt_kept"], out0["pt_kept"])
    for key in ("cam_P", "cam_f", "cam_d", "pts"):
        a, b = out1[key].astype(np.float64), out0[key].astype(np.float64)
        assert np.abs(a - b).max() <= 2e-6 * max(np.abs(b).max(), 1e-30), key
    assert len(log1) == len(log0)
    for l0, l1 in zip(log0, log1):
        t0, t1 = l0.split(), l1.split()
        assert t0[0] == t1[0] and t0[-4:] == t1[-4:]


@pytest.mark.parametrize("robust", [1, 0])
def test_batch_blocks_match_oracle(robust):
    sizes = [(8, 300), (20, 900), (5, 200), (33, 1500), (12, 500), (3, 120)]
    blocks = _blocks(500 + 10 * robust, sizes)
    flags = capi.make_flags(robust=robust, solver=1)
    print(f"[batch robust={robust}] {len(blocks)} blocks", flush=True)
    res = api.calc_plsa_batch(blocks, flags)
    for i, (sc, got) in enumerate(zip(blocks, res)):
        _compare_block(sc, flags, got, f"block {i}")


def test_batch_graph_replay_equals_eager_and_reset():
    blocks = _blocks(700, [(10, 400), (25, 1200), (6, 250), (40, 2000)])
    flags = capi.make_flags(robust=1, solver=1)
    b = api.Batch(blocks, flags)
    print("[batch graph] tuning:", b.tuning(), flush=True)
    rc, n, fin = b.run(-1)
    assert rc == 0 and fin
    rd, g1 = b.download()
    assert "graph launches 0" not in b.tuning()
    b.reset()
    b.set_graph(False)
    rc, n2, fin = b.run(-1)
    assert rc == 0 and fin and n2 == n
    rd, e1 = b.download()
    b.reset()
    b.set_graph(True)
    rc, n3, fin = b.run(-1)
    rd, g2 = b.download()
    b.close()
    for (s1, o1, t1, _), (s2, o2, t2, _), (s3, o3, t3, _) in zip(g1, e1, g2):
        assert s1 == s2 == s3 == 0
        for key in ("cam_P", "cam_f", "cam_d", "pts", "obs_kept", "pt_kept"):
            assert np.array_equal(o1[key], o2[key]) and np.array_equal(o1[key], o3[key]), key
        assert np.array_equal(t1["trace_sig0"], t2["trace_sig0"])
        assert np.array_equal(t1["trace_sig0"], t3["trace_sig0"])


def test_batch_block_error_is_isolated():
    """A block whose redundancy is negative stops with the reference's error
    (rbacore.cpp:468-476) and keeps its initial state; the other blocks solve
    as if alone."""
    good = _blocks(900, [(9, 350), (14, 600)])
    bad = scenes.generate(None, seed=950, n_cams=5, n_pts=8, track="uniform2_8", outlier_frac=0.0)
    blocks = [good[0], bad, good[1]]
    flags = capi.make_flags(robust=1, solver=1)
    res = api.calc_plsa_batch(blocks, flags)
    assert res[1][0] == capi.RPBA_E_LOW_REDUNDANCY
    assert any("Low redundancy" in line for line in res[1][3])
    for i in (0, 2):
        _compare_block(blocks[i], flags, res[i], f"block {i}")
    _compare_block(bad, flags, res[1], "failing block")


@pytest.mark.parametrize("solver", [1, 0])
def test_batch_fp32_records_match_single_block_solves(solver):
    """precision = 1 in the batch: each block equals the single-block GPU
    solve of the same form (FP32 records; dense solve, or the PCG segmented
    by block) under the trajectory rule (the two differ only in the order of
    the cost sums, and with the PCG in the order of its dot products)."""
    blocks = _blocks(1200, [(10, 400), (22, 900), (6, 250)])
    flags = capi.make_flags(robust=1, solver=solver, precision=1)
    res = api.calc_plsa_batch(blocks, flags)
    for i, (sc, (rc1, out1, st1, log1)) in enumerate(zip(blocks, res)):
        rc0, out0, st0, log0 = api.calc_plsa(sc, flags)
        print(f"[batch fp32 block {i}] itt {st0['itt']} {st1['itt']}", flush=True)
        assert rc1 == rc0 == 0 and st1["itt"] == st0["itt"]
        check_trajectory(st1["trace_sig0"], st0["trace_sig0"], st1["trace_pcg"], st0["trace_pcg"],
                         st1["trace_accept"], st0["trace_accept"], st1["trace_nrobs"], st0["trace_nrobs"])
        assert np.array_equal(out1["obs_kept"], out0["obs_kept"])


def _compare_block_pcg(sc, flags, got, tag):
    """The PCG form of a block against the oracle's SparsePCG solve of it: the
    bars of the single-block golden tests (trajectory rule, converged sigma
    within 1e-6, labels bit for bit, point mean and focal lengths)."""
    rc1, out1, st1, log1 = got
    rc0, out0, st0, log0 = oracle.solve(sc, flags, fsites=0)
    print(f"[{tag}] n_cams={sc['n_cams']} itt oracle {st0['itt']} batch {st1['itt']} pcg oracle "
          f"{list(st0['trace_pcg'])} batch {list(st1['trace_pcg'])}", flush=True)
    assert rc1 == rc0
    assert st1["passes_done"] == st0["passes_done"] and st1["itt"] == st0["itt"]
    check_trajectory(st1["trace_sig0"], st0["trace_sig0"], st1["trace_pcg"], st0["trace_pcg"],
                     st1["trace_accept"], st0["trace_accept"], st1["trace_nrobs"], st0["trace_nrobs"])
    assert st1["pcg_iters"] == int(np.sum(st1["trace_pcg"])) and st1["pcg_iters"] > 0
    for p in range(st0["passes_done"]):
        assert abs(st1["sigma_out"][p] - st0["sigma_out"][p]) <= 1e-6 * abs(st0["sigma_out"][p])
        for key in ("redundancy", "nobs", "nunknowns"):
            assert st1[key][p] == st0[key][p]
    assert np.array_equal(out1["obs_kept"], out0["obs_kept"])
    assert np.array_equal(out1["pt_kept"], out0["pt_kept"])
    assert np.allclose(out1["pts"].astype(np.float64).mean(axis=0), out0["pts"].astype(np.float64).mean(axis=0),
                       rtol=1e-5, atol=1e-6)
    assert np.allclose(out1["cam_f"], out0["cam_f"], rtol=1e-6)
    assert len(log1) == len(log0)


@pytest.mark.parametrize("robust", [1, 0])
def test_batch_pcg_blocks_match_oracle(robust):
    """solver 0: one SpMV over the union per PCG iteration, each block with its
    own alpha, beta, best iterate, stopping rule and cap; blocks that stop
    early (and blocks that sit out the rest of a pass) leave the others
    running.  Each block against the oracle's PCG solve of it alone."""
    sizes = [(8, 300), (20, 900), (5, 200), (33, 1500), (12, 500), (3, 120)]
    blocks = _blocks(600 + 10 * robust, sizes)
    flags = capi.make_flags(robust=robust, solver=0)
    print(f"[batch pcg robust={robust}] {len(blocks)} blocks", flush=True)
    res = api.calc_plsa_batch(blocks, flags)
    for i, (sc, got) in enumerate(zip(blocks, res)):
        _compare_block_pcg(sc, flags, got, f"pcg block {i}")


def test_batch_pcg_graph_replay_equals_eager_and_reset():
    """The three graphs of the PCG form (set-up, four PCG iterations, step)
    replayed against eager launches, bitwise, and a reset re-run."""
    blocks = _blocks(760, [(10, 400), (25, 1200), (6, 250), (40, 2000)])
    flags = capi.make_flags(robust=1, solver=0)
    b = api.Batch(blocks, flags)
    print("[batch pcg graph] tuning:", b.tuning(), flush=True)
    rc, n, fin = b.run(-1)
    assert rc == 0 and fin
    rd, g1 = b.download()
    assert "segmented" in b.tuning() and "graph launches 0" not in b.tuning()
    b.reset()
    b.set_graph(False)
    rc, n2, fin = b.run(-1)
    assert rc == 0 and fin and n2 == n
    rd, e1 = b.download()
    b.reset()
    b.set_graph(True)
    rc, n3, fin = b.run(-1)
    rd, g2 = b.download()
    b.close()
    for (s1, o1, t1, _), (s2, o2, t2, _), (s3, o3, t3, _) in zip(g1, e1, g2):
        assert s1 == s2 == s3 == 0
        for key in ("cam_P", "cam_f", "cam_d", "pts", "obs_kept", "pt_kept"):
            assert np.array_equal(o1[key], o2[key]) and np.array_equal(o1[key], o3[key]), key
        for key in ("trace_sig0", "trace_pcg"):
            assert np.array_equal(t1[key], t2[key]) and np.array_equal(t1[key], t3[key]), key


def test_batch_pcg_matches_single_block_solves():
    """Each block of the segmented PCG against the single-block GPU solve of
    the same block (solver 0), under the trajectory rule."""
    blocks = _blocks(1300, [(10, 400), (22, 900), (6, 250), (30, 1400)])
    flags = capi.make_flags(robust=1, solver=0)
    res = api.calc_plsa_batch(blocks, flags)
    for i, (sc, (rc1, out1, st1, log1)) in enumerate(zip(blocks, res)):
        rc0, out0, st0, log0 = api.calc_plsa(sc, flags)
        print(f"[batch pcg vs single {i}] itt {st0['itt']} {st1['itt']} pcg {list(st0['trace_pcg'])} "
              f"{list(st1['trace_pcg'])}", flush=True)
        assert rc1 == rc0 == 0 and st1["itt"] == st0["itt"]
        check_trajectory(st1["trace_sig0"], st0["trace_sig0"], st1["trace_pcg"], st0["trace_pcg"],
                         st1["trace_accept"], st0["trace_accept"], st1["trace_nrobs"], st0["trace_nrobs"])
        assert np.array_equal(out1["obs_kept"], out0["obs_kept"])


def test_batch_pcg_spmv_bytes_accounting():
    """rpba_batch_spmv_bytes over a timed solve: with one SpMV launch per PCG
    iteration of the slowest running block, its bytes lie between the
    per-camera terms alone (152 + 648 bytes a camera per launch) and the
    whole union S read on every launch; with every block still running on
    the first launch, more than the per-camera terms."""
    blocks = _blocks(1400, [(10, 400), (22, 900), (6, 250), (30, 1400)])
    b = api.Batch(blocks, capi.make_flags(robust=1, solver=0))
    b.set_timing(True)
    rc, n, fin = b.run(1)
    sz = b.sizes()  # the union S of pass 0 (RobustDelete keeps every block)
    rc, n, fin = b.run(-1)
    assert rc == 0 and fin
    spmv_bytes = b.spmv_bytes()
    ms, launches = b.kernel_time("spmv")
    b.close()
    print(f"[batch spmv bytes] {spmv_bytes:.4g} bytes over {launches} launches, {ms:.3f} ms", flush=True)
    assert launches > 0 and ms > 0
    per_cam = 800.0 * sz["cams"]
    assert spmv_bytes > per_cam * launches
    assert spmv_bytes <= (648.0 * sz["blocks"] + per_cam) * launches
