"""ctypes binding of libneurohip.so (include/neuro_hip.h).

Loads the in-tree library and raises if it is missing: the product path has
no fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libneurohip.so")

NH_OK, NH_ERR_ARG, NH_ERR_HIP, NH_ERR_NO_DEVICE, NH_ERR_WINDOW, NH_ERR_NONE, \
    NH_ERR_UNSUPPORTED, NH_ERR_CAPACITY, NH_ERR_NET = range(9)


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


class Config(C.Structure):
    _fields_ = [
        ("dt_ms", C.c_float), ("synapse_tau_ms", C.c_float), ("readout_tau_ms", C.c_float),
        ("noise_rate_hz", C.c_float), ("noise_kick_mv", C.c_float), ("noise_sd", C.c_float),
        ("seed", C.c_uint64), ("record_spikes", C.c_int32),
        ("spike_record_steps", C.c_uint32), ("int_weights", C.c_int32),
        ("int_scale", C.c_float), ("raster_steps", C.c_uint32),
        ("exec_mode", C.c_int32), ("step_shape", C.c_int32), ("graph_steps", C.c_uint32),
        ("block", C.c_uint32), ("lds_bins", C.c_uint32),
    ]


class Ne2Info(C.Structure):
    _fields_ = [
        ("version", C.c_uint32), ("dt_ms", C.c_float), ("conduction_speed", C.c_float),
        ("synapse_tau_ms", C.c_float), ("readout_tau_ms", C.c_float),
        ("noise_rate_hz", C.c_float), ("noise_kick_mv", C.c_float), ("seed", C.c_uint64),
        ("n_types", C.c_uint32), ("n_groups", C.c_uint32), ("n_inputs", C.c_uint32),
        ("n_samples", C.c_uint32), ("n_readouts", C.c_uint32), ("n_scripts", C.c_uint32),
    ]


class NHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[nh status {code}] {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"HIP backend library missing: {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    vp, i32, u32, u64, f32, f64 = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_float, C.c_double
    P = C.POINTER
    sig = {
        "nh_config_default": (None, [P(Config)]),
        "nh_last_error": (C.c_char_p, []),
        "nh_available": (i32, []),
        "nh_check_net": (i32, [P(NetDesc), P(Config)]),
        "nh_create": (i32, [P(NetDesc), P(Config), P(vp)]),
        "nh_destroy": (None, [vp]),
        "nh_run_steps": (i32, [vp, u64]),
        "nh_run_ms": (i32, [vp, f32]),
        "nh_reset": (i32, [vp]),
        "nh_sync": (i32, [vp]),
        "nh_n": (u32, [vp]),
        "nh_step_count": (u64, [vp]),
        "nh_time_ms": (f64, [vp]),
        "nh_dev_mem": (u64, [vp]),
        "nh_read_trace": (i32, [vp, vp, u32]),
        "nh_read_state": (i32, [vp, u32, vp, u32]),
        "nh_read_syn_current": (i32, [vp, vp, u32]),
        "nh_spikes_total": (i32, [vp, P(u64)]),
        "nh_mean_rate_hz": (i32, [vp, P(f64)]),
        "nh_read_pn_total": (i32, [vp, vp, u32]),
        "nh_events_total": (i32, [vp, P(u64)]),
        "nh_mark_counters": (i32, [vp]),
        "nh_counts_since_mark": (i32, [vp, P(u64), P(u64)]),
        "nh_spike_row": (i32, [vp, u64, vp, u32]),
        "nh_last_spikes": (i32, [vp, vp, u32, P(u32)]),
        "nh_spike_counts": (i32, [vp, u64, vp, u32]),
        "nh_spike_counts_noalloc": (i32, [vp, u64, vp, u32]),
        "nh_set_input_constant": (i32, [vp, u32, f32]),
        "nh_input_values": (i32, [vp, vp, u32]),
        "nh_play_input": (i32, [vp, u32, vp, u64]),
        "nh_loop_input": (i32, [vp, u32, vp, u64]),
        "nh_set_external": (i32, [vp, u32, f32]),
        "nh_set_external_all": (i32, [vp, vp, u32]),
        "nh_read_external": (i32, [vp, vp, u32]),
        "nh_set_noise": (i32, [vp, f32, f32]),
        "nh_set_noise_sd": (i32, [vp, f32]),
        "nh_set_v_watch": (i32, [vp, vp, u32]),
        "nh_drain_watch": (i32, [vp, vp, u32, P(u32)]),
        "nh_set_grid_watches": (i32, [vp, u32, vp, vp]),
        "nh_grid_traces": (i32, [vp, f32, vp, u64]),
        "nh_read_spikes": (i32, [vp, P(u64), vp, u64, vp, u64, P(u64), P(u64)]),
        "nh_last_run_ms": (i32, [vp, P(f64)]),
        "nh_last_run_tail": (i32, [vp, P(f64), P(u64)]),
        "nh_profile_steps": (i32, [vp, u64, P(f64), P(u64), P(f64), P(u64)]),
        "nh_autotune": (i32, [vp, u32, C.c_char_p, u32]),
        "nh_launch_shape": (i32, [vp, C.POINTER(C.c_int32)]),
        "nh_test_pair_wait": (i32, [vp, u32, u32, P(u64)]),
        "nh_probe_copy": (i32, [u64, u32, u32, P(f64)]),
        "nh_probe_launch": (i32, [u32, u32, P(f64)]),
        "nh_batch_autotune": (i32, [vp, u32, C.c_char_p, u32]),
        "nh_net_free": (None, [vp]),
        "nh_net_view": (i32, [vp, P(NetDesc)]),
        "nh_net_load_ne2": (i32, [C.c_char_p, u64, C.c_float, P(vp), P(Ne2Info)]),
        "nh_net_type_name": (i32, [vp, u32, C.c_char_p, u32]),
        "nh_net_neuron_types": (i32, [vp, vp, u32]),
        "nh_net_input_name": (i32, [vp, u32, C.c_char_p, u32]),
        "nh_net_group": (i32, [vp, u32, C.c_char_p, u32, vp, u32, P(u32)]),
        "nh_net_bench": (i32, [u32, u32, u64, f32, f32, f32, P(vp)]),
        "nh_net_cube": (i32, [u32, f64, u64, f32, P(vp)]),
        "nh_net_grid": (i32, [u32, u32, u32, f64, f64, f64, f32, f32, C.c_int, u32, u32, f32,
                              u64, f32, P(vp)]),
        "nh_rng_stream_new": (i32, [u64, u32, f32, P(vp)]),
        "nh_rng_stream_free": (None, [vp]),
        "nh_rng_stream_window": (i32, [vp, u32, C.c_int, vp, u32, f32, vp]),
        "nh_batch_create": (i32, [u32, vp, vp, P(vp)]),
        "nh_batch_destroy": (None, [vp]),
        "nh_batch_run_steps": (i32, [vp, u64]),
        "nh_batch_sync": (i32, [vp]),
        "nh_batch_spikes_total": (i32, [vp, u32, P(u64)]),
        "nh_batch_events_total": (i32, [vp, P(u64)]),
        "nh_batch_mark_counters": (i32, [vp]),
        "nh_batch_counts_since_mark": (i32, [vp, P(u64), P(u64)]),
        "nh_batch_read_state": (i32, [vp, u32, u32, vp, u32]),
        "nh_batch_spike_counts": (i32, [vp, u32, u64, vp, u32]),
        "nh_batch_read_spikes": (i32, [vp, u32, P(u64), vp, u64, vp, u64, P(u64), P(u64)]),
        "nh_batch_play_input": (i32, [vp, u32, u32, vp, u64]),
        "nh_batch_last_run_ms": (i32, [vp, P(f64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != NH_OK:
        msg = lib().nh_last_error()
        raise NHError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Address of a contiguous numpy array (None for empty)."""
    if a is None or a.size == 0:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def as_arr(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def make_desc(net):
    """NetDesc over the arrays of a network dict; returns (desc, keepalive)."""
    n = int(net["n"])
    keep = {
        "src_ptr": as_arr(net["src_ptr"], np.uint32), "tgt": as_arr(net["tgt"], np.uint32),
        "w": as_arr(net["w"], np.float32), "delay": as_arr(net["delay"], np.uint16),
        "model_ids": as_arr(net["model_ids"], np.uint8), "model": as_arr(net["model"], np.uint8),
        "p": as_arr(net["p"], np.float32), "x0": as_arr(net["x0"], np.float32),
    }
    tau = net.get("tau")
    keep["tau"] = as_arr(tau, np.float32) if tau is not None and len(tau) else None
    ni = int(net.get("n_inputs", 0) or 0)
    if ni and net.get("stim_ptr") is not None:
        keep["stim_ptr"] = as_arr(net["stim_ptr"], np.uint32)
        keep["stim_input"] = as_arr(net["stim_input"], np.uint32)
        keep["stim_w"] = as_arr(net["stim_w"], np.float32)
    else:
        keep["stim_ptr"] = keep["stim_input"] = keep["stim_w"] = None
    d = NetDesc()
    d.n = n
    d.n_syn = len(keep["tgt"])
    for f in ("src_ptr", "tgt", "w", "delay", "tau", "model_ids", "model", "p", "x0",
              "stim_ptr", "stim_input", "stim_w"):
        setattr(d, f, ptr(keep[f]))
    d.n_models = len(keep["model_ids"])
    d.n_params = int(net["n_params"])
    d.n_state = int(net["n_state"])
    d.n_inputs = ni
    return d, keep
