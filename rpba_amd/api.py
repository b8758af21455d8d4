"""Python mirror of the reference's solve entry point over the C ABI.

``calc_plsa`` follows ``CalcPLSA`` (rpba.h:693-703, rpbacore.cpp:28-531): it
takes a block (cameras, points, observations in file order), overwrites the
cameras and points with the adjusted values, marks the observations and
points that survive the robust deletion, and returns the status together
with the log lines the reference writes to its ``os`` stream.
"""
import ctypes as C

from . import capi, lib


def calc_plsa(scene, flags=None, trace_cap=512):
    """Returns (status, outputs, stats, log lines)."""
    flags = flags or capi.make_flags()
    blk = capi.Block(scene)
    tr = capi.Trace(trace_cap)
    logc = capi.LogCollector()
    rc = lib().rpba_calc_plsa(C.byref(blk.c), C.byref(flags), logc.fn, None, C.byref(tr.stats))
    return rc, blk.result(), tr.as_dict(), logc.lines


def partition_images(scene, nparts):
    """The image partition of the partitioned solve (rpba_partition_images):
    one part id per camera.  Host code only; needs no GPU."""
    import numpy as np
    cam = np.ascontiguousarray(scene["obs_cam"], dtype=np.int32)
    pt = np.ascontiguousarray(scene["obs_pt"], dtype=np.int32)
    part = np.zeros(scene["n_cams"], dtype=np.int32)
    P = C.POINTER(C.c_int32)
    rc = lib().rpba_partition_images(scene["n_cams"], len(cam), cam.ctypes.data_as(P), pt.ctypes.data_as(P),
                                     scene["n_pts"], nparts, part.ctypes.data_as(P))
    if rc != 0:
        raise RuntimeError(f"rpba_partition_images failed with status {rc}")
    return part


def dense_solve_batch(systems, rhs, variant=-1, cap=0):
    """Solve the SPD systems (list of (9k x 9k) arrays) with the batched dense
    Cholesky of the dense form; returns the list of solutions."""
    import numpy as np
    ncam = np.array([s.shape[0] // 9 for s in systems], dtype=np.int32)
    S = np.concatenate([np.ascontiguousarray(s, dtype=np.float64).ravel() for s in systems])
    b = np.concatenate([np.asarray(r, dtype=np.float64) for r in rhs])
    x = np.zeros_like(b)
    P = C.POINTER
    rc = lib().rpba_dense_solve_batch(len(systems), ncam.ctypes.data_as(P(C.c_int32)), S.ctypes.data_as(P(C.c_double)),
                                      b.ctypes.data_as(P(C.c_double)), x.ctypes.data_as(P(C.c_double)), variant, cap)
    if rc != 0:
        raise RuntimeError(f"rpba_dense_solve_batch failed with status {rc}")
    out, o = [], 0
    for s in systems:
        out.append(x[o:o + s.shape[0]])
        o += s.shape[0]
    return out


class Context:
    """Resumable solve (rpba_ctx_*): create uploads the block, run advances
    the LM iterations, download writes the stored state back."""

    def __init__(self, scene, flags=None):
        self.flags = flags or capi.make_flags()
        self.blk = capi.Block(scene)
        self.logc = capi.LogCollector()
        self.h = C.c_void_p()
        rc = lib().rpba_ctx_create(C.byref(self.h), C.byref(self.blk.c), C.byref(self.flags), self.logc.fn, None)
        if rc != 0:
            raise RuntimeError(f"rpba_ctx_create failed with status {rc}")

    def run(self, max_iters=-1):
        done, fin = C.c_int32(), C.c_int32()
        rc = lib().rpba_ctx_run(self.h, max_iters, C.byref(done), C.byref(fin))
        return rc, done.value, bool(fin.value)

    def reset(self):
        return lib().rpba_ctx_reset(self.h)

    def set_timing(self, on):
        return lib().rpba_ctx_set_timing(self.h, 1 if on else 0)

    def set_pcg_budget(self, n):
        """Every PCG solve stops at n iterations and the run then reports
        RPBA_E_BUDGET (0: the reference's cap)."""
        return lib().rpba_ctx_set_pcg_budget(self.h, n)

    def set_graph(self, on):
        """PCG iterations replayed from a hipGraph, or launched one by one (the default)."""
        return lib().rpba_ctx_set_graph(self.h, 1 if on else 0)

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int64()
        lib().rpba_ctx_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def progress(self):
        """(pass, itt, stored sig0) of the running solve."""
        ps, it, sg = C.c_int32(), C.c_int32(), C.c_double()
        lib().rpba_ctx_progress(self.h, C.byref(ps), C.byref(it), C.byref(sg))
        return ps.value, it.value, sg.value

    def sizes(self):
        out = (C.c_int64 * 6)()
        lib().rpba_ctx_sizes(self.h, out)
        return dict(zip(["cams", "pts", "obs", "pairs", "blocks", "sdim"], list(out)))

    def set_lp(self, model):
        """Attach a graph network (LpModel) for preconditioner = 2."""
        self._lp = model
        return lib().rpba_ctx_set_lp(self.h, model.h if model is not None else None)

    def last_omega(self):
        w = C.c_float()
        lib().rpba_ctx_last_omega(self.h, C.byref(w))
        return w.value

    def lp_sample(self, omegas):
        """Camera graph of the last LM iteration's S and the PCG iterations of
        that S and b for each weight in omegas."""
        import numpy as np
        P = C.POINTER
        nc = self.blk.n_cams
        om = np.ascontiguousarray(omegas, dtype=np.float32)
        its = np.zeros(om.size, dtype=np.int32)
        feat = np.zeros((nc, capi.LP_FEATURES), dtype=np.float32)
        off = np.zeros(nc + 1, dtype=np.int32)
        ne = C.c_int64(0)
        rc = lib().rpba_ctx_lp_sample(self.h, 0, None, None, None, None, None, None, C.byref(ne))
        if rc != 0:
            raise RuntimeError(f"rpba_ctx_lp_sample failed with status {rc}")
        nbr = np.zeros(max(ne.value, 1), dtype=np.int32)
        w = np.zeros(max(ne.value, 1), dtype=np.float32)
        rc = lib().rpba_ctx_lp_sample(self.h, om.size, om.ctypes.data_as(P(C.c_float)), its.ctypes.data_as(P(C.c_int32)),
                                      feat.ctypes.data_as(P(C.c_float)), off.ctypes.data_as(P(C.c_int32)),
                                      nbr.ctypes.data_as(P(C.c_int32)), w.ctypes.data_as(P(C.c_float)), C.byref(ne))
        if rc != 0:
            raise RuntimeError(f"rpba_ctx_lp_sample failed with status {rc}")
        return {"feat": feat, "adj_off": off, "adj_nbr": nbr[:ne.value], "adj_w": w[:ne.value], "iters": its}

    def tuning(self):
        """The dense form's autotuner choice and timings (one line)."""
        buf = C.create_string_buffer(512)
        lib().rpba_ctx_tuning(self.h, buf, 512)
        return buf.value.decode()

    def download(self, trace_cap=512):
        tr = capi.Trace(trace_cap)
        rc = lib().rpba_ctx_download(self.h, C.byref(tr.stats))
        return rc, self.blk.result(), tr.as_dict()

    def close(self):
        if self.h:
            lib().rpba_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched independent blocks (rpba_batch_*): all blocks' LM iterations
    run together on the GPU; download writes every block back."""

    def __init__(self, scenes_, flags=None):
        self.flags = flags or capi.make_flags()
        self.blks = [capi.Block(s) for s in scenes_]
        self.arr = (capi.RpbaBlock * len(self.blks))(*[b.c for b in self.blks])
        self.lines = [[] for _ in self.blks]
        self.fn = capi.BATCH_LOG_FN(self._cb)
        self.h = C.c_void_p()
        rc = lib().rpba_batch_create(C.byref(self.h), len(self.blks), self.arr, C.byref(self.flags), self.fn, None)
        if rc != 0:
            raise RuntimeError(f"rpba_batch_create failed with status {rc}")

    def _cb(self, block, line, user):
        self.lines[block].append(line.decode())

    def run(self, max_iters=-1):
        done, fin = C.c_int32(), C.c_int32()
        rc = lib().rpba_batch_run(self.h, max_iters, C.byref(done), C.byref(fin))
        return rc, done.value, bool(fin.value)

    def reset(self):
        return lib().rpba_batch_reset(self.h)

    def set_graph(self, on):
        return lib().rpba_batch_set_graph(self.h, 1 if on else 0)

    def set_timing(self, on):
        return lib().rpba_batch_set_timing(self.h, 1 if on else 0)

    def kernel_time(self, name):
        ms, n = C.c_double(), C.c_int64()
        lib().rpba_batch_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def spmv_bytes(self):
        """solver 0: algorithmic bytes of the SpMV launches timed so far"""
        v = C.c_double()
        lib().rpba_batch_spmv_bytes(self.h, C.byref(v))
        return v.value

    def tuning(self):
        buf = C.create_string_buffer(1024)
        lib().rpba_batch_tuning(self.h, buf, 1024)
        return buf.value.decode()

    def sizes(self):
        out = (C.c_int64 * 6)()
        lib().rpba_batch_sizes(self.h, out)
        return dict(zip(["cams", "pts", "obs", "pairs", "blocks", "dense_doubles"], list(out)))

    def download(self, trace_cap=512):
        n = len(self.blks)
        trs = [capi.Trace(trace_cap) for _ in range(n)]
        stats = (capi.RpbaStats * n)(*[t.stats for t in trs])
        status = (C.c_int32 * n)()
        rc = lib().rpba_batch_download(self.h, stats, status)
        out = []
        for i in range(n):
            trs[i].stats = stats[i]
            out.append((int(status[i]), self.blks[i].result(), trs[i].as_dict(), list(self.lines[i])))
        return rc, out

    def close(self):
        if self.h:
            lib().rpba_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calc_plsa_batch(scenes_, flags=None, trace_cap=512):
    """Returns a list of (status, outputs, stats, log lines), one per block."""
    b = Batch(scenes_, flags)
    rc, n, fin = b.run(-1)
    rd, out = b.download(trace_cap)
    b.close()
    return out


def lp_graphs(graphs, labels=None):
    """Pack a list of graphs (dicts feat, adj_off, adj_nbr, adj_w) into one
    rpba_lp_graphs; returns (struct, arrays kept alive)."""
    import numpy as np
    P = C.POINTER
    n = [g["feat"].shape[0] for g in graphs]
    goff = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
    feat = np.ascontiguousarray(np.concatenate([g["feat"] for g in graphs]), dtype=np.float32)
    offs, nbrs, ws = [np.zeros(1, np.int64)], [], []
    e0 = 0
    for g, n0 in zip(graphs, goff[:-1]):
        offs.append(np.asarray(g["adj_off"][1:], np.int64) + e0)
        nbrs.append(np.asarray(g["adj_nbr"], np.int64) + n0)
        ws.append(np.asarray(g["adj_w"], np.float32))
        e0 += len(g["adj_nbr"])
    adj_off = np.concatenate(offs).astype(np.int32)
    adj_nbr = np.ascontiguousarray(np.concatenate(nbrs) if nbrs else np.zeros(0), dtype=np.int32)
    adj_w = np.ascontiguousarray(np.concatenate(ws) if ws else np.zeros(0), dtype=np.float32)
    lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.float32)
    keep = [goff, feat, adj_off, adj_nbr, adj_w, lab]
    st = capi.RpbaLpGraphs(len(graphs), int(goff[-1]), int(adj_nbr.size),
                           goff.ctypes.data_as(P(C.c_int32)), feat.ctypes.data_as(P(C.c_float)),
                           adj_off.ctypes.data_as(P(C.c_int32)),
                           adj_nbr.ctypes.data_as(P(C.c_int32)) if adj_nbr.size else None,
                           adj_w.ctypes.data_as(P(C.c_float)) if adj_w.size else None,
                           lab.ctypes.data_as(P(C.c_float)) if lab is not None else None)
    return st, keep


class LpModel:
    """The learned preconditioner's graph network on the GPU (rpba_lp_*)."""

    def __init__(self, params=None, seed=0):
        import numpy as np
        self.h = C.c_void_p()
        p = None
        if params is not None:
            self._p = np.ascontiguousarray(params, dtype=np.float32)
            p = self._p.ctypes.data_as(C.POINTER(C.c_float))
        rc = lib().rpba_lp_create(C.byref(self.h), p, seed)
        if rc != 0:
            raise RuntimeError(f"rpba_lp_create failed with status {rc}")

    @staticmethod
    def n_params():
        return lib().rpba_lp_n_params()

    @classmethod
    def load(cls, path):
        import numpy as np
        return cls(np.fromfile(path, dtype="<f4"))

    def params(self):
        import numpy as np
        out = np.zeros(self.n_params(), dtype=np.float32)
        lib().rpba_lp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def save(self, path):
        self.params().astype("<f4").tofile(path)

    def predict(self, graphs):
        import numpy as np
        st, keep = lp_graphs(graphs)
        out = np.zeros(len(graphs), dtype=np.float32)
        rc = lib().rpba_lp_predict(self.h, C.byref(st), out.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            raise RuntimeError(f"rpba_lp_predict failed with status {rc}")
        return out

    def gradient(self, graphs, labels):
        import numpy as np
        st, keep = lp_graphs(graphs, labels)
        om = np.zeros(len(graphs), dtype=np.float32)
        g = np.zeros(self.n_params(), dtype=np.float32)
        loss = C.c_float()
        P = C.POINTER
        rc = lib().rpba_lp_gradient(self.h, C.byref(st), om.ctypes.data_as(P(C.c_float)), C.byref(loss),
                                    g.ctypes.data_as(P(C.c_float)))
        if rc != 0:
            raise RuntimeError(f"rpba_lp_gradient failed with status {rc}")
        return om, loss.value, g

    def train(self, graphs, labels, steps, lr):
        import numpy as np
        st, keep = lp_graphs(graphs, labels)
        tr = np.zeros(max(steps, 1), dtype=np.float32)
        rc = lib().rpba_lp_train(self.h, C.byref(st), steps, lr, tr.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            raise RuntimeError(f"rpba_lp_train failed with status {rc}")
        return tr[:steps]

    def close(self):
        if self.h:
            lib().rpba_lp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
