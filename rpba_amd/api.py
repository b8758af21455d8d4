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
