"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, always as the checker.  The product
library never links or calls it.

The C sources restate the reference CPU backend (sim/src/cpu.rs) and the
net-format builders; see oracle.h for the file:line map.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class NetDesc(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("n_syn", C.c_uint64),
        ("src_ptr", C.c_void_p), ("tgt", C.c_void_p), ("w", C.c_void_p),
        ("delay", C.c_void_p), ("tau", C.c_void_p),
        ("n_models", C.c_uint32), ("model_ids", C.c_void_p), ("model", C.c_void_p),
        ("n_params", C.c_uint32), ("n_state", C.c_uint32),
        ("p", C.c_void_p), ("x0", C.c_void_p),
        ("n_inputs", C.c_uint32), ("stim_ptr", C.c_void_p),
        ("stim_input", C.c_void_p), ("stim_w", C.c_void_p),
    ]


class Cfg(C.Structure):
    _fields_ = [
        ("dt_ms", C.c_float), ("synapse_tau_ms", C.c_float), ("readout_tau_ms", C.c_float),
        ("noise_rate_hz", C.c_float), ("noise_kick_mv", C.c_float), ("noise_sd", C.c_float),
        ("seed", C.c_uint64), ("record_spikes", C.c_int32),
        ("spike_record_steps", C.c_uint32), ("int_weights", C.c_int32),
        ("int_scale", C.c_float), ("exp_mode", C.c_int32), ("precision", C.c_int32),
    ]


class RngState(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


def lib():
    """Load liboracle.so (built by `make -C oracle`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle not built: {_LIB_PATH} (run make -C oracle)")
    L = C.CDLL(_LIB_PATH)
    vp, i32, u32, u64, f32, f64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64, C.c_float, C.c_double
    sig = {
        "or_cfg_default": (None, [C.POINTER(Cfg)]),
        "or_sim_new": (vp, [C.POINTER(NetDesc), C.POINTER(Cfg), C.c_char_p, C.c_int]),
        "or_sim_free": (None, [vp]),
        "or_sim_reset": (None, [vp]),
        "or_sim_run": (C.c_int, [vp, u64, vp, vp, u64, C.POINTER(C.c_uint64)]),
        "or_sim_t": (u64, [vp]),
        "or_sim_spikes_total": (u64, [vp]),
        "or_sim_ring_depth": (u32, [vp]),
        "or_sim_events_total": (u64, [vp]),
        "or_sim_state": (None, [vp, u32, f64p]),
        "or_sim_syn_current": (None, [vp, f64p]),
        "or_sim_trace": (None, [vp, f64p]),
        "or_sim_pn_total": (None, [vp, u64p]),
        "or_sim_n_xsyn": (u64, [vp]),
        "or_sim_xsyn": (None, [vp, f64p, u64]),
        "or_sim_spike_row": (C.c_int, [vp, u64, u64p]),
        "or_sim_spike_counts": (C.c_int, [vp, u64, u32p]),
        "or_sim_set_external": (None, [vp, u32, f32]),
        "or_sim_set_input_constant": (None, [vp, u32, f32]),
        "or_sim_play_input": (None, [vp, u32, f32p, u64, C.c_int]),
        "or_sim_set_noise": (None, [vp, f32, f32]),
        "or_sim_set_grid_watches": (None, [vp, u32, u32p, u32p]),
        "or_sim_grid_traces": (None, [vp, f32, f32p]),
        "or_noise_hash": (u32, [u32, u64, u64]),
        "or_noise_fires": (C.c_int, [u32, u64, u64, f32]),
        "or_gauss_draw": (f32, [u32, u64, u64]),
        "or_portable_expf": (f32, [f32]),
        "or_ring_put": (i32, [f32, f32]),
        "or_rng_init": (None, [C.POINTER(RngState), u64]),
        "or_rng_next_u64": (u64, [C.POINTER(RngState)]),
        "or_rng_f64": (f64, [C.POINTER(RngState)]),
        "or_rng_f32": (f32, [C.POINTER(RngState)]),
        "or_rng_uniform": (f64, [C.POINTER(RngState), f64, f64]),
        "or_rng_below": (u64, [C.POINTER(RngState), u64]),
        "or_rng_gauss": (f64, [C.POINTER(RngState), f64, f64]),
        "or_netb_new": (vp, []),
        "or_netb_free": (None, [vp]),
        "or_netb_set_speed": (None, [vp, f32]),
        "or_netb_add_neuron": (u32, [vp, f32, f32, f32, C.c_int]),
        "or_netb_set_kind": (None, [vp, u32, C.c_int]),
        "or_netb_set_model": (None, [vp, u32, C.c_int, f32p]),
        "or_netb_connect": (u64, [vp, u32, u32, f32]),
        "or_netb_push_synapse": (None, [vp, u32, u32, f32, f32, f32]),
        "or_netb_grid2d": (u32, [vp, u32, u32, f32, f32, f32, f32]),
        "or_netb_grid3d": (u32, [vp, u32, u32, u32, f32, f32, f32, f32]),
        "or_netb_create_connections": (None, [vp, u32, f64, f64, f32, C.POINTER(RngState)]),
        "or_netb_set_exc_inh": (None, [vp, u32, f64, C.POINTER(RngState)]),
        "or_netb_set_exc_inh_exact": (None, [vp, u32, u64, C.POINTER(RngState)]),
        "or_netb_set_synapse_tau_dist": (None, [vp, u32, u32, f64, f64, C.POINTER(RngState)]),
        "or_netb_bench_net": (None, [vp, u32, u32, u64, f32, f32]),
        "or_netb_add_input": (u32, [vp, u32, u32p, f32p]),
        "or_netb_n": (u32, [vp]),
        "or_netb_n_syn": (u64, [vp]),
        "or_netb_csr": (C.c_int, [vp, f32, u32p, u32p, f32p, u16p, f32p, C.c_char_p, C.c_int]),
        "or_netb_params": (None, [vp, C.POINTER(C.c_uint32), u8p, u8p, C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32), f32p, f32p]),
        "or_netb_n_inputs": (u32, [vp]),
        "or_netb_n_stim": (u64, [vp]),
        "or_netb_stim_map": (None, [vp, u32p, u32p, f32p]),
        "or_netb_positions": (None, [vp, f32p]),
        "or_netb_kinds": (None, [vp, u8p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


# ---------------------------------------------------------------- arithmetic
def noise_hash(i, t, seed):
    return lib().or_noise_hash(i, t, seed)


def noise_fires(i, t, seed, p):
    return bool(lib().or_noise_fires(i, t, seed, p))


def gauss_draw(i, t, seed):
    return lib().or_gauss_draw(i, t, seed)


def portable_expf(x):
    return lib().or_portable_expf(x)


def ring_put(w, scale):
    return lib().or_ring_put(w, scale)


class Rng:
    """net-format/src/rng.rs:17-81 restated."""

    def __init__(self, seed):
        self.st = RngState()
        lib().or_rng_init(C.byref(self.st), seed)

    def next_u64(self):
        return lib().or_rng_next_u64(C.byref(self.st))

    def f64(self):
        return lib().or_rng_f64(C.byref(self.st))

    def f32(self):
        return lib().or_rng_f32(C.byref(self.st))

    def uniform(self, lo, hi):
        return lib().or_rng_uniform(C.byref(self.st), lo, hi)

    def below(self, n):
        return lib().or_rng_below(C.byref(self.st), n)

    def gauss(self, mean, sd):
        return lib().or_rng_gauss(C.byref(self.st), mean, sd)


class NetBuilder:
    """Editable network restated from net-format (network.rs, builders.rs)."""

    EXC, INH, REGULAR = 0, 1, 2

    def __init__(self):
        self.h = lib().or_netb_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_netb_free(self.h)
            self.h = None

    def set_speed(self, v):
        lib().or_netb_set_speed(self.h, v)

    def add_neuron(self, pos, kind=0):
        return lib().or_netb_add_neuron(self.h, pos[0], pos[1], pos[2], kind)

    def set_kind(self, i, kind):
        lib().or_netb_set_kind(self.h, i, kind)

    def set_model(self, i, model, params):
        lib().or_netb_set_model(self.h, i, model, np.ascontiguousarray(params, np.float32))

    def connect(self, src, tgt, w):
        return lib().or_netb_connect(self.h, src, tgt, w)

    def push_synapse(self, src, tgt, w, delay_ms, tau_ms=0.0):
        lib().or_netb_push_synapse(self.h, src, tgt, w, delay_ms, tau_ms)

    def grid2d(self, w, h, dist, center=(0.0, 0.0, 0.0)):
        return lib().or_netb_grid2d(self.h, w, h, dist, *center)

    def grid3d(self, w, h, d, dist, center=(0.0, 0.0, 0.0)):
        return lib().or_netb_grid3d(self.h, w, h, d, dist, *center)

    def create_connections(self, group, scaling, lam, weight, rng):
        lib().or_netb_create_connections(self.h, group, scaling, lam, weight, C.byref(rng.st))

    def set_exc_inh(self, group, frac, rng):
        lib().or_netb_set_exc_inh(self.h, group, frac, C.byref(rng.st))

    def set_synapse_tau_dist(self, src_group, tgt_group, mean, sd, rng):
        lib().or_netb_set_synapse_tau_dist(self.h, src_group, tgt_group, mean, sd, C.byref(rng.st))

    def set_exc_inh_exact(self, group, n_inh, rng):
        lib().or_netb_set_exc_inh_exact(self.h, group, n_inh, C.byref(rng.st))

    def bench_net(self, n, k, seed, w_scale, w_inh):
        lib().or_netb_bench_net(self.h, n, k, seed, w_scale, w_inh)

    def add_input(self, targets):
        tg = np.ascontiguousarray([t for t, _ in targets], np.uint32)
        wt = np.ascontiguousarray([w for _, w in targets], np.float32)
        return lib().or_netb_add_input(self.h, len(tg), tg, wt)

    @property
    def n(self):
        return lib().or_netb_n(self.h)

    def positions(self):
        out = np.zeros(3 * max(self.n, 1), np.float32)
        lib().or_netb_positions(self.h, out)
        return out[: 3 * self.n].reshape(-1, 3)

    def kinds(self):
        out = np.zeros(max(self.n, 1), np.uint8)
        lib().or_netb_kinds(self.h, out)
        return out[: self.n]

    def build(self, dt_ms):
        """Flatten to the simulator layout (synapse_csr + NeuronParams +
        StimMap).  Raises ValueError with the reference's message on a
        delay past the u16 ring."""
        L = lib()
        n = self.n
        E = L.or_netb_n_syn(self.h)
        src_ptr = np.zeros(n + 1, np.uint32)
        tgt = np.zeros(max(E, 1), np.uint32)
        w = np.zeros(max(E, 1), np.float32)
        delay = np.zeros(max(E, 1), np.uint16)
        tau = np.zeros(max(E, 1), np.float32)
        err = C.create_string_buffer(512)
        if L.or_netb_csr(self.h, dt_ms, src_ptr, tgt, w, delay, tau, err, 512) != 0:
            raise ValueError(err.value.decode())
        nm, npar, nst = C.c_uint32(), C.c_uint32(), C.c_uint32()
        model_ids = np.zeros(8, np.uint8)
        model = np.zeros(max(n, 1), np.uint8)
        p = np.zeros(max(n, 1) * 10, np.float32)
        x0 = np.zeros(max(n, 1) * 2, np.float32)
        L.or_netb_params(self.h, C.byref(nm), model_ids, model, C.byref(npar), C.byref(nst), p, x0)
        ni = L.or_netb_n_inputs(self.h)
        ns = L.or_netb_n_stim(self.h)
        stim_ptr = np.zeros(n + 1, np.uint32)
        stim_input = np.zeros(max(ns, 1), np.uint32)
        stim_w = np.zeros(max(ns, 1), np.float32)
        L.or_netb_stim_map(self.h, stim_ptr, stim_input, stim_w)
        npar, nst = npar.value, max(nst.value, 2)
        return dict(
            n=n, src_ptr=src_ptr, tgt=tgt[:E], w=w[:E], delay=delay[:E], tau=tau[:E],
            model_ids=model_ids[: nm.value].copy(), model=model[:n].copy(),
            n_params=npar, n_state=nst, p=p[: n * npar].copy(), x0=x0[: n * nst].copy(),
            n_inputs=ni, stim_ptr=stim_ptr, stim_input=stim_input[:ns].copy(),
            stim_w=stim_w[:ns].copy(), kind=self.kinds().copy(),
        )


def default_cfg(**kw):
    c = Cfg()
    lib().or_cfg_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def _arr(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class OracleSim:
    """CpuSim (sim/src/cpu.rs) restated; f32 or f64 by cfg.precision."""

    def __init__(self, net, cfg=None, **kw):
        L = lib()
        cfg = cfg if cfg is not None else default_cfg()
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.cfg = cfg
        self.n = int(net["n"])
        self._keep = {
            "src_ptr": _arr(net["src_ptr"], np.uint32), "tgt": _arr(net["tgt"], np.uint32),
            "w": _arr(net["w"], np.float32), "delay": _arr(net["delay"], np.uint16),
            "tau": _arr(net["tau"] if net.get("tau") is not None else np.zeros(len(net["tgt"])),
                        np.float32),
            "model_ids": _arr(net["model_ids"], np.uint8), "model": _arr(net["model"], np.uint8),
            "p": _arr(net["p"], np.float32), "x0": _arr(net["x0"], np.float32),
            "stim_ptr": _arr(net.get("stim_ptr", np.zeros(self.n + 1)), np.uint32),
            "stim_input": _arr(net.get("stim_input", np.zeros(0)), np.uint32),
            "stim_w": _arr(net.get("stim_w", np.zeros(0)), np.float32),
        }
        k = self._keep
        d = NetDesc()
        d.n = self.n
        d.n_syn = len(k["tgt"])
        for f in ("src_ptr", "tgt", "w", "delay", "tau", "model_ids", "model", "p", "x0",
                  "stim_ptr", "stim_input", "stim_w"):
            a = k[f]
            setattr(d, f, a.ctypes.data if a.size else None)
        d.n_models = len(k["model_ids"])
        d.n_params = int(net["n_params"])
        d.n_state = int(net["n_state"])
        d.n_inputs = int(net.get("n_inputs", 0))
        self.n_state = d.n_state
        self.n_inputs = d.n_inputs
        self._desc = d
        err = C.create_string_buffer(512)
        self.h = L.or_sim_new(C.byref(d), C.byref(cfg), err, 512)
        if not self.h:
            raise ValueError(err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_sim_free(self.h)
            self.h = None

    # ---- Backend-like surface ----
    @property
    def t(self):
        return lib().or_sim_t(self.h)

    step_count = t

    def ring_depth(self):
        return lib().or_sim_ring_depth(self.h)

    def run_steps(self, steps, record=False):
        """Run; with record=True returns (counts[steps], ids) -- the raster
        as per-step counts and ascending ids per step (recorded in chunks
        whose id buffer holds the worst case, every neuron every step)."""
        if not record:
            lib().or_sim_run(self.h, steps, None, None, 0, None)
            return None
        cnt = np.zeros(max(steps, 1), np.uint32)
        chunk = max(1, (1 << 22) // max(self.n, 1))
        parts, done = [], 0
        while done < steps:
            k = min(chunk, steps - done)
            cap = max(self.n * k, 1)
            ids = np.zeros(cap, np.uint32)
            nid = C.c_uint64()
            rc = lib().or_sim_run(self.h, k, cnt[done:].ctypes.data, ids.ctypes.data, cap,
                                  C.byref(nid))
            assert rc == 0
            parts.append(ids[: nid.value])
            done += k
        return cnt[:steps], (np.concatenate(parts) if parts else np.zeros(0, np.uint32))

    def run_ms(self, ms):
        self.run_steps(int(round(ms / self.cfg.dt_ms)))

    def reset_state(self):
        lib().or_sim_reset(self.h)

    def state(self, slot):
        out = np.zeros(self.n, np.float64)
        if slot < self.n_state:
            lib().or_sim_state(self.h, slot, out)
        return out

    def v(self):
        return self.state(0)

    def syn_current(self):
        out = np.zeros(self.n, np.float64)
        lib().or_sim_syn_current(self.h, out)
        return out

    def trace(self):
        out = np.zeros(self.n, np.float64)
        lib().or_sim_trace(self.h, out)
        return out

    def pn_total(self):
        out = np.zeros(self.n, np.uint64)
        lib().or_sim_pn_total(self.h, out)
        return out

    def xsyn(self):
        m = lib().or_sim_n_xsyn(self.h)
        out = np.zeros(max(m, 1), np.float64)
        lib().or_sim_xsyn(self.h, out, m)
        return out[:m]

    def spikes_total(self):
        return lib().or_sim_spikes_total(self.h)

    def events_total(self):
        return lib().or_sim_events_total(self.h)

    def spike_row(self, back):
        out = np.zeros((self.n + 63) // 64 or 1, np.uint64)
        ok = lib().or_sim_spike_row(self.h, back, out)
        return out[: (self.n + 63) // 64] if ok else None

    def spike_counts(self, rows):
        out = np.zeros(max(self.n, 1), np.uint32)
        rc = lib().or_sim_spike_counts(self.h, min(rows, (1 << 64) - 1), out)
        if rc != 0:
            raise ValueError(f"spike_counts window {rows} exceeds the recorded steps")
        return out[: self.n]

    def set_external(self, i, amps):
        lib().or_sim_set_external(self.h, i, amps)

    def set_input_constant(self, k, v):
        lib().or_sim_set_input_constant(self.h, k, v)

    def play_input(self, k, series, looped=False):
        s = np.ascontiguousarray(series, np.float32)
        lib().or_sim_play_input(self.h, k, s if s.size else np.zeros(1, np.float32), s.size,
                                int(looped))

    def loop_input(self, k, series):
        self.play_input(k, series, looped=True)

    def set_noise(self, rate_hz, kick_mv):
        lib().or_sim_set_noise(self.h, rate_hz, kick_mv)

    def set_grid_watches(self, groups):
        sizes = np.ascontiguousarray([len(g) for g in groups] or [0], np.uint32)
        mem = np.ascontiguousarray([m for g in groups for m in g] or [0], np.uint32)
        lib().or_sim_set_grid_watches(self.h, len(groups), sizes, mem)
        self._gsizes = [len(g) for g in groups]

    def grid_traces(self, tau_ms):
        tot = sum(self._gsizes)
        out = np.zeros(max(tot, 1), np.float32)
        lib().or_sim_grid_traces(self.h, tau_ms, out)
        res, o = [], 0
        for s in self._gsizes:
            res.append(out[o:o + s].copy())
            o += s
        return res
