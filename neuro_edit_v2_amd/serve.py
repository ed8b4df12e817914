"""Served form: continuous batching of LSM trial requests over one resident
arena (nh_server_* in include/neuro_hip.h).

``Server.submit`` hands a request to the worker thread of the library; the
request is admitted when neurons, synapses and a slot are free, runs its
steps together with every other in-flight request, and ``Server.wait``
returns its per-window spike counts and timing stamps.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class ServerConfig(C.Structure):
    _fields_ = [
        ("cap_neurons", C.c_uint64), ("cap_synapses", C.c_uint64), ("cap_stim", C.c_uint64),
        ("max_requests", C.c_uint32), ("max_delay", C.c_uint32), ("n_inputs", C.c_uint32),
        ("n_windows", C.c_uint32), ("chunk_steps", C.c_uint32), ("int_weights", C.c_int32),
        ("int_scale", C.c_float), ("hbm_fraction", C.c_float),
    ]


class RequestStats(C.Structure):
    _fields_ = [("submit_ns", C.c_uint64), ("admit_ns", C.c_uint64), ("done_ns", C.c_uint64),
                ("t0", C.c_uint64), ("spikes", C.c_uint64), ("n", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("events", C.c_uint64), ("spikes", C.c_uint64),
                ("admitted", C.c_uint64), ("completed", C.c_uint64), ("peak_active", C.c_uint64),
                ("cap_neurons", C.c_uint64), ("cap_synapses", C.c_uint64), ("gpu_ms", C.c_double),
                ("paired_rounds", C.c_uint64)]


def _sig():
    lib = L.lib()
    if getattr(lib, "_srv_sig", False):
        return lib
    vp, u32, u64, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
    P = C.POINTER
    for name, res, args in [
        ("nh_server_config_default", None, [P(ServerConfig)]),
        ("nh_server_create", C.c_int, [P(ServerConfig), P(vp)]),
        ("nh_server_destroy", None, [vp]),
        ("nh_server_submit", C.c_int, [vp, P(L.NetDesc), P(L.Config), vp, u64, u64, u32, P(u64)]),
        ("nh_server_wait", C.c_int, [vp, u64, u32, vp, u64, P(RequestStats)]),
        ("nh_server_get_stats", C.c_int, [vp, P(ServerStats)]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._srv_sig = True
    return lib


class Server:
    def __init__(self, **kw):
        lib = _sig()
        c = ServerConfig()
        lib.nh_server_config_default(C.byref(c))
        for k, v in kw.items():
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
        self.cfg = c
        h = C.c_void_p()
        L.check(lib.nh_server_create(C.byref(c), C.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            _sig().nh_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, net, cfg, stim, total_steps, win_steps):
        desc, keep = L.make_desc(net)
        st = np.ascontiguousarray(stim, np.float32)
        rid = C.c_uint64()
        L.check(_sig().nh_server_submit(self._h, C.byref(desc), C.byref(cfg.to_c()), L.ptr(st),
                                        st.shape[0] if st.size else 0, total_steps, win_steps,
                                        C.byref(rid)))
        self._n[rid.value] = int(net["n"])
        return rid.value

    def wait(self, rid, timeout_ms=60000):
        """(counts[n_windows, n], stats dict)."""
        n = self._n[rid]
        nw = self.cfg.n_windows
        out = np.zeros(max(nw * n, 1), np.uint32)
        st = RequestStats()
        L.check(_sig().nh_server_wait(self._h, rid, timeout_ms, L.ptr(out), nw * n, C.byref(st)))
        del self._n[rid]
        return out[: nw * n].reshape(nw, n), {
            "submit_ns": st.submit_ns, "admit_ns": st.admit_ns, "done_ns": st.done_ns,
            "t0": st.t0, "spikes": st.spikes, "n": st.n}

    def stats(self):
        s = ServerStats()
        L.check(_sig().nh_server_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}


def window_counts(counts_per_step, ids, n, win_steps, n_windows):
    """Per-window spike counts from a raster (counts per step, ids)."""
    out = np.zeros((n_windows, n), np.uint32)
    off = np.concatenate([[0], np.cumsum(counts_per_step.astype(np.int64))])
    for w in range(n_windows):
        a, b = off[w * win_steps], off[min((w + 1) * win_steps, len(counts_per_step))]
        np.add.at(out[w], ids[a:b], 1)
    return out
