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
