"""``Sim`` -- the reference ``Backend`` trait (sim/src/lib.rs:119-184) over
the gfx950 HIP path, and ``SimConfig`` (sim/src/state.rs:7-49).

Method names and meanings follow the trait; the spike-count window error
of ``spike_counts`` (sim/src/cpu.rs:463-468) raises ``ValueError`` and
``spike_row`` returns ``None`` outside the recorded window, as the trait's
``Result``/``Option`` do.
"""
import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib as L


def available():
    """cuda_available (sim/src/lib.rs:192-203) for the HIP backend."""
    return bool(L.lib().nh_available())


@dataclass
class SimConfig:
    """SimConfig -- defaults as SimConfig::default (sim/src/state.rs:35-49)
    plus this backend's additions (Gaussian sd, int scale, raster stream,
    execution shape)."""
    dt_ms: float = 0.1
    synapse_tau_ms: float = 2.0
    readout_tau_ms: float = 2.0
    noise_rate_hz: float = 0.0
    noise_kick_mv: float = 20.0
    noise_sd: float = 0.0
    seed: int = 1
    record_spikes: bool = False
    spike_record_steps: int = 512
    int_weights: bool = False
    int_scale: float = 4096.0
    raster_steps: int = 0
    exec_mode: int = 0
    step_shape: int = -1
    graph_steps: int = 0
    block: int = 0
    lds_bins: int = 0

    def to_c(self):
        c = L.Config()
        for f in fields(self):
            v = getattr(self, f.name)
            setattr(c, f.name, int(v) if isinstance(v, bool) else v)
        return c


class Sim:
    """One network on the GPU (CudaSim's role, sim/src/cuda.rs)."""

    def __init__(self, net, cfg=None, **kw):
        cfg = cfg if cfg is not None else SimConfig()
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.cfg = cfg
        lib = L.lib()
        desc, keep = L.make_desc(net)
        h = C.c_void_p()
        L.check(lib.nh_create(C.byref(desc), C.byref(cfg.to_c()), C.byref(h)))
        self._h = h
        self.n = int(net["n"])
        self.n_inputs = int(net.get("n_inputs", 0) or 0)
        self._groups = []

    def close(self):
        if getattr(self, "_h", None):
            L.lib().nh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- Backend trait ----
    def kind(self):
        return "hip"

    def step_count(self):
        return L.lib().nh_step_count(self._h)

    def time_ms(self):
        return L.lib().nh_time_ms(self._h)

    def run_steps(self, n):
        L.check(L.lib().nh_run_steps(self._h, int(n)))

    def run_ms(self, ms):
        L.check(L.lib().nh_run_ms(self._h, float(ms)))

    def reset_state(self):
        L.check(L.lib().nh_reset(self._h))

    def synchronize(self):
        L.check(L.lib().nh_sync(self._h))

    def dev_mem_bytes(self):
        return L.lib().nh_dev_mem(self._h)

    def _f32(self, fn, *args):
        out = np.zeros(max(self.n, 1), np.float32)
        L.check(fn(self._h, *args, L.ptr(out), self.n))
        return out[: self.n]

    def trace(self):
        return self._f32(L.lib().nh_read_trace)

    def state(self, slot):
        if slot >= 2:
            return np.zeros(0, np.float32)
        return self._f32(L.lib().nh_read_state, slot)

    def v(self):
        return self.state(0)

    def syn_current(self):
        return self._f32(L.lib().nh_read_syn_current)

    def spikes_total(self):
        v = C.c_uint64()
        L.check(L.lib().nh_spikes_total(self._h, C.byref(v)))
        return v.value

    def events_total(self):
        v = C.c_uint64()
        L.check(L.lib().nh_events_total(self._h, C.byref(v)))
        return v.value

    def mark_counters(self):
        """Snapshot the per-neuron spike totals on the device, in stream
        order and without waiting (measurement helper)."""
        L.check(L.lib().nh_mark_counters(self._h))

    def counts_since_mark(self):
        """(spikes, synaptic events) since the last mark_counters."""
        sp, ev = C.c_uint64(), C.c_uint64()
        L.check(L.lib().nh_counts_since_mark(self._h, C.byref(sp), C.byref(ev)))
        return sp.value, ev.value

    def mean_rate_hz(self):
        v = C.c_double()
        L.check(L.lib().nh_mean_rate_hz(self._h, C.byref(v)))
        return v.value

    def pn_total(self):
        out = np.zeros(max(self.n, 1), np.uint64)
        L.check(L.lib().nh_read_pn_total(self._h, L.ptr(out), self.n))
        return out[: self.n]

    def spike_row(self, back):
        w = (self.n + 63) // 64
        out = np.zeros(max(w, 1), np.uint64)
        rc = L.lib().nh_spike_row(self._h, int(back), L.ptr(out), w)
        if rc == L.NH_ERR_NONE:
            return None
        L.check(rc)
        return out[:w]

    def last_spikes(self):
        out = np.zeros(max(self.n, 1), np.uint32)
        k = C.c_uint32()
        L.check(L.lib().nh_last_spikes(self._h, L.ptr(out), self.n, C.byref(k)))
        return out[: k.value]

    def spike_counts(self, rows, no_alloc=False):
        """Backend::spike_counts; no_alloc counts into the library's
        scratch instead of a device buffer of the sim's own."""
        out = np.zeros(max(self.n, 1), np.uint32)
        rows = min(int(rows), (1 << 64) - 1)
        fn = L.lib().nh_spike_counts_noalloc if no_alloc else L.lib().nh_spike_counts
        rc = fn(self._h, rows, L.ptr(out), self.n)
        if rc == L.NH_ERR_WINDOW:
            raise ValueError(L.lib().nh_last_error().decode())
        L.check(rc)
        return out[: self.n]

    def set_input_constant(self, k, value):
        L.check(L.lib().nh_set_input_constant(self._h, k, value))

    def input_values(self):
        out = np.zeros(max(self.n_inputs, 1), np.float32)
        L.check(L.lib().nh_input_values(self._h, L.ptr(out), self.n_inputs))
        return out[: self.n_inputs]

    def play_input(self, k, series):
        s = np.ascontiguousarray(series, np.float32)
        L.check(L.lib().nh_play_input(self._h, k, L.ptr(s), s.size))

    def loop_input(self, k, series):
        s = np.ascontiguousarray(series, np.float32)
        L.check(L.lib().nh_loop_input(self._h, k, L.ptr(s), s.size))

    def set_external(self, i, amps):
        L.check(L.lib().nh_set_external(self._h, int(i), float(amps)))

    def set_external_all(self, ext):
        e = np.ascontiguousarray(ext, np.float32)
        L.check(L.lib().nh_set_external_all(self._h, L.ptr(e), e.size))

    def external(self):
        return self._f32(L.lib().nh_read_external)

    def set_noise(self, rate_hz, kick_mv):
        L.check(L.lib().nh_set_noise(self._h, rate_hz, kick_mv))

    def set_noise_sd(self, sd):
        L.check(L.lib().nh_set_noise_sd(self._h, sd))

    def set_v_watch(self, ids):
        a = np.ascontiguousarray(ids, np.uint32)
        L.check(L.lib().nh_set_v_watch(self._h, L.ptr(a), a.size))
        self._nwatch = a.size

    def drain_watch(self):
        """[state slot][watch][samples oldest-first] (lists of arrays)."""
        nw = getattr(self, "_nwatch", 0)
        cap = 4096
        out = np.zeros(max(2 * nw * cap, 1), np.float32)
        k = C.c_uint32()
        L.check(L.lib().nh_drain_watch(self._h, L.ptr(out), cap, C.byref(k)))
        o = out[: 2 * nw * cap].reshape(2, nw, cap) if nw else out[:0].reshape(2, 0, cap)
        return [[o[s, c, : k.value].copy() for c in range(nw)] for s in range(2)]

    def set_grid_watches(self, groups):
        sizes = np.ascontiguousarray([len(g) for g in groups], np.uint32)
        mem = np.ascontiguousarray([m for g in groups for m in g], np.uint32)
        L.check(L.lib().nh_set_grid_watches(self._h, len(groups), L.ptr(sizes), L.ptr(mem)))
        self._groups = [len(g) for g in groups]

    def grid_traces(self, tau_ms):
        tot = sum(self._groups)
        out = np.zeros(max(tot, 1), np.float32)
        L.check(L.lib().nh_grid_traces(self._h, tau_ms, L.ptr(out), tot))
        res, o = [], 0
        for s in self._groups:
            res.append(out[o:o + s].copy())
            o += s
        return res

    def read_spikes(self):
        """Drain the raster: (step0, counts[steps], ids) -- ids ascending per step."""
        lib = L.lib()
        s0, ns, ni = C.c_uint64(), C.c_uint64(), C.c_uint64()
        rc = lib.nh_read_spikes(self._h, C.byref(s0), None, 0, None, 0, C.byref(ns), C.byref(ni))
        if rc not in (L.NH_OK, L.NH_ERR_CAPACITY):
            L.check(rc)
        cnt = np.zeros(max(ns.value, 1), np.uint32)
        ids = np.zeros(max(ni.value, 1), np.uint32)
        L.check(lib.nh_read_spikes(self._h, C.byref(s0), L.ptr(cnt), ns.value, L.ptr(ids),
                                   ni.value, C.byref(ns), C.byref(ni)))
        return s0.value, cnt[: ns.value], ids[: ni.value]

    def profile_steps(self, n):
        """Run n steps with an event pair around every step-kernel launch;
        returns (main_ms, main_launches, scatter_ms, scatter_launches)."""
        a, b = C.c_double(), C.c_double()
        na, nb = C.c_uint64(), C.c_uint64()
        L.check(L.lib().nh_profile_steps(self._h, int(n), C.byref(a), C.byref(na), C.byref(b),
                                         C.byref(nb)))
        return a.value, na.value, b.value, nb.value

    def autotune(self, trial_steps=256):
        """Pick the fastest launch shape / workgroup / graph length from
        step 0 (the sim is reset afterwards); returns the trial report."""
        buf = C.create_string_buffer(4096)
        L.check(L.lib().nh_autotune(self._h, trial_steps, buf, 4096))
        return buf.value.decode()

    def launch_shape(self):
        """Launch shape of the next run: 0 two kernels per step, 1 one fused
        kernel per step, 2 two steps per launch."""
        v = C.c_int32()
        L.check(L.lib().nh_launch_shape(self._h, C.byref(v)))
        return v.value

    def last_run_ms(self):
        v = C.c_double()
        L.check(L.lib().nh_last_run_ms(self._h, C.byref(v)))
        return v.value

    def last_run_tail(self):
        """(ms, launches): the last run's time after its first step launch
        and the step launches in it (back-to-back time per launch)."""
        v, n = C.c_double(), C.c_uint64()
        L.check(L.lib().nh_last_run_tail(self._h, C.byref(v), C.byref(n)))
        return v.value, int(n.value)
