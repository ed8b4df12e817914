"""Batched form: many independent networks of uneven size packed by offset
tables into one arena and stepped by the same kernel in one launch per
step (nh_batch_* in include/neuro_hip.h).

Each network keeps its own seed, noise, stimulus and raster; its result is
the plain path's result for it alone (int ring: bit-exact).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .sim import SimConfig


class Batch:
    def __init__(self, nets, cfgs):
        if isinstance(cfgs, SimConfig):
            cfgs = [cfgs] * len(nets)
        self.n = [int(x["n"]) for x in nets]
        self.n_inputs = [int(x.get("n_inputs", 0) or 0) for x in nets]
        descs = (L.NetDesc * len(nets))()
        self._keep = []
        for k, net in enumerate(nets):
            d, keep = L.make_desc(net)
            descs[k] = d
            self._keep.append(keep)
        cc = (L.Config * len(nets))()
        for k, c in enumerate(cfgs):
            cc[k] = c.to_c()
        h = C.c_void_p()
        L.check(L.lib().nh_batch_create(len(nets), C.cast(descs, C.c_void_p),
                                        C.cast(cc, C.c_void_p), C.byref(h)))
        self._h = h
        self.cfgs = cfgs

    def close(self):
        if getattr(self, "_h", None):
            L.lib().nh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run_steps(self, n):
        L.check(L.lib().nh_batch_run_steps(self._h, int(n)))

    def synchronize(self):
        L.check(L.lib().nh_batch_sync(self._h))

    def play_input(self, net, k, series):
        s = np.ascontiguousarray(series, np.float32)
        L.check(L.lib().nh_batch_play_input(self._h, net, k, L.ptr(s), s.size))

    def spikes_total(self, net):
        v = C.c_uint64()
        L.check(L.lib().nh_batch_spikes_total(self._h, net, C.byref(v)))
        return v.value

    def events_total(self):
        v = C.c_uint64()
        L.check(L.lib().nh_batch_events_total(self._h, C.byref(v)))
        return v.value

    def mark_counters(self):
        """Snapshot the spike totals on the device, without waiting."""
        L.check(L.lib().nh_batch_mark_counters(self._h))

    def counts_since_mark(self):
        """(spikes, synaptic events) of the whole batch since the last mark."""
        sp, ev = C.c_uint64(), C.c_uint64()
        L.check(L.lib().nh_batch_counts_since_mark(self._h, C.byref(sp), C.byref(ev)))
        return sp.value, ev.value

    def state(self, net, slot):
        out = np.zeros(max(self.n[net], 1), np.float32)
        L.check(L.lib().nh_batch_read_state(self._h, net, slot, L.ptr(out), self.n[net]))
        return out[: self.n[net]]

    def spike_counts(self, net, rows):
        """Per-neuron spike counts of network `net` over its last `rows`
        steps (Backend::spike_counts); ValueError for a window between the
        record ring and t."""
        out = np.zeros(max(self.n[net], 1), np.uint32)
        rows = min(int(rows), (1 << 64) - 1)
        rc = L.lib().nh_batch_spike_counts(self._h, net, rows, L.ptr(out), self.n[net])
        if rc == L.NH_ERR_WINDOW:
            raise ValueError(L.lib().nh_last_error().decode())
        L.check(rc)
        return out[: self.n[net]]

    def read_spikes(self, net):
        lib = L.lib()
        s0, ns, ni = C.c_uint64(), C.c_uint64(), C.c_uint64()
        rc = lib.nh_batch_read_spikes(self._h, net, C.byref(s0), None, 0, None, 0, C.byref(ns),
                                      C.byref(ni))
        if rc not in (L.NH_OK, L.NH_ERR_CAPACITY):
            L.check(rc)
        cnt = np.zeros(max(ns.value, 1), np.uint32)
        ids = np.zeros(max(ni.value, 1), np.uint32)
        L.check(lib.nh_batch_read_spikes(self._h, net, C.byref(s0), L.ptr(cnt), ns.value,
                                         L.ptr(ids), ni.value, C.byref(ns), C.byref(ni)))
        return s0.value, cnt[: ns.value], ids[: ni.value]

    def autotune(self, trial_steps=256):
        buf = C.create_string_buffer(4096)
        L.check(L.lib().nh_batch_autotune(self._h, trial_steps, buf, 4096))
        return buf.value.decode()

    def last_run_ms(self):
        v = C.c_double()
        L.check(L.lib().nh_batch_last_run_ms(self._h, C.byref(v)))
        return v.value
