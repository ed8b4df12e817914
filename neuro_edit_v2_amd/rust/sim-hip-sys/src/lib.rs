//! Raw FFI declarations of `include/neuro_hip.h` (libneurohip, the gfx950
//! backend of the reference `sim` crate).  One `extern "C"` item per C entry
//! point and one `#[repr(C)]` struct per C struct, field for field and in
//! order, so every layout equals the C one.  The safe wrapper that
//! implements `sim::Backend` on top of these is `sim/src/hip.rs`.
//!
//! Conventions of the library: every `i32` result is an `nh_status`
//! (`NH_OK` = 0); the message of the last failure on the calling thread is
//! `nh_last_error()`.  Outputs are caller-allocated with explicit lengths.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int};

pub const NH_OK: c_int = 0;
pub const NH_ERR_ARG: c_int = 1;
pub const NH_ERR_HIP: c_int = 2;
pub const NH_ERR_NO_DEVICE: c_int = 3;
pub const NH_ERR_WINDOW: c_int = 4;
pub const NH_ERR_NONE: c_int = 5;
pub const NH_ERR_UNSUPPORTED: c_int = 6;
pub const NH_ERR_CAPACITY: c_int = 7;
pub const NH_ERR_NET: c_int = 8;

/// `nh_net_desc`: the flat layouts the reference builds (`SynCsr`,
/// `NeuronParams`, `StimMap`) as plain pointers.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct NhNetDesc {
    pub n: u32,
    pub n_syn: u64,
    pub src_ptr: *const u32,
    pub tgt: *const u32,
    pub w: *const f32,
    pub delay: *const u16,
    pub tau: *const f32,
    pub n_models: u32,
    pub model_ids: *const u8,
    pub model: *const u8,
    pub n_params: u32,
    pub n_state: u32,
    pub p: *const f32,
    pub x0: *const f32,
    pub n_inputs: u32,
    pub stim_ptr: *const u32,
    pub stim_input: *const u32,
    pub stim_w: *const f32,
}

/// `nh_config`: `SimConfig` plus the backend's additions.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct NhConfig {
    pub dt_ms: f32,
    pub synapse_tau_ms: f32,
    pub readout_tau_ms: f32,
    pub noise_rate_hz: f32,
    pub noise_kick_mv: f32,
    pub noise_sd: f32,
    pub seed: u64,
    pub record_spikes: i32,
    pub spike_record_steps: u32,
    pub int_weights: i32,
    pub int_scale: f32,
    pub raster_steps: u32,
    pub exec_mode: i32,
    pub step_shape: i32,
    pub graph_steps: u32,
    pub block: u32,
    pub lds_bins: u32,
}

/// `nh_ne2_info`: the `Meta` and table sizes of a loaded `.ne2` file.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NhNe2Info {
    pub version: u32,
    pub dt_ms: f32,
    pub conduction_speed: f32,
    pub synapse_tau_ms: f32,
    pub readout_tau_ms: f32,
    pub noise_rate_hz: f32,
    pub noise_kick_mv: f32,
    pub seed: u64,
    pub n_types: u32,
    pub n_groups: u32,
    pub n_inputs: u32,
    pub n_samples: u32,
    pub n_readouts: u32,
    pub n_scripts: u32,
}

/// `nh_server_config`
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NhServerConfig {
    pub cap_neurons: u64,
    pub cap_synapses: u64,
    pub cap_stim: u64,
    pub max_requests: u32,
    pub max_delay: u32,
    pub n_inputs: u32,
    pub n_windows: u32,
    pub chunk_steps: u32,
    pub int_weights: i32,
    pub int_scale: f32,
    pub hbm_fraction: f32,
}

/// `nh_request_stats`
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NhRequestStats {
    pub submit_ns: u64,
    pub admit_ns: u64,
    pub done_ns: u64,
    pub t0: u64,
    pub spikes: u64,
    pub n: u32,
}

/// `nh_server_stats`
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct NhServerStats {
    pub steps: u64,
    pub events: u64,
    pub spikes: u64,
    pub admitted: u64,
    pub completed: u64,
    pub peak_active: u64,
    pub cap_neurons: u64,
    pub cap_synapses: u64,
    pub gpu_ms: f64,
    pub paired_rounds: u64,
}

/// Opaque handles (`nh_sim`, `nh_net`, `nh_batch`, `nh_server`,
/// `nh_rng_stream`).
#[repr(C)]
pub struct NhSim {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NhNet {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NhBatch {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NhServer {
    _private: [u8; 0],
}
#[repr(C)]
pub struct NhRngStream {
    _private: [u8; 0],
}

extern "C" {
    // ---- configuration, errors, device probe
    pub fn nh_config_default(cfg: *mut NhConfig);
    pub fn nh_last_error() -> *const c_char;
    pub fn nh_available() -> c_int;

    // ---- one network (Backend)
    pub fn nh_check_net(net: *const NhNetDesc, cfg: *const NhConfig) -> c_int;
    pub fn nh_create(net: *const NhNetDesc, cfg: *const NhConfig, out: *mut *mut NhSim) -> c_int;
    pub fn nh_destroy(s: *mut NhSim);
    pub fn nh_run_steps(s: *mut NhSim, n_steps: u64) -> c_int;
    pub fn nh_run_ms(s: *mut NhSim, ms: f32) -> c_int;
    pub fn nh_reset(s: *mut NhSim) -> c_int;
    pub fn nh_sync(s: *mut NhSim) -> c_int;
    pub fn nh_n(s: *const NhSim) -> u32;
    pub fn nh_step_count(s: *const NhSim) -> u64;
    pub fn nh_time_ms(s: *const NhSim) -> f64;
    pub fn nh_dev_mem(s: *const NhSim) -> u64;
    pub fn nh_read_trace(s: *mut NhSim, out: *mut f32, n: u32) -> c_int;
    pub fn nh_read_state(s: *mut NhSim, slot: u32, out: *mut f32, n: u32) -> c_int;
    pub fn nh_read_syn_current(s: *mut NhSim, out: *mut f32, n: u32) -> c_int;
    pub fn nh_spikes_total(s: *mut NhSim, out: *mut u64) -> c_int;
    pub fn nh_mean_rate_hz(s: *mut NhSim, out: *mut f64) -> c_int;
    pub fn nh_read_pn_total(s: *mut NhSim, out: *mut u64, n: u32) -> c_int;
    pub fn nh_events_total(s: *mut NhSim, out: *mut u64) -> c_int;
    pub fn nh_mark_counters(s: *mut NhSim) -> c_int;
    pub fn nh_counts_since_mark(s: *mut NhSim, spikes: *mut u64, events: *mut u64) -> c_int;
    pub fn nh_spike_row(s: *mut NhSim, back: u64, words: *mut u64, n_words: u32) -> c_int;
    pub fn nh_last_spikes(s: *mut NhSim, ids: *mut u32, cap: u32, n_out: *mut u32) -> c_int;
    pub fn nh_spike_counts(s: *mut NhSim, rows: u64, out: *mut u32, n: u32) -> c_int;
    pub fn nh_spike_counts_noalloc(s: *mut NhSim, rows: u64, out: *mut u32, n: u32) -> c_int;
    pub fn nh_set_input_constant(s: *mut NhSim, input: u32, value: f32) -> c_int;
    pub fn nh_input_values(s: *mut NhSim, out: *mut f32, n_inputs: u32) -> c_int;
    pub fn nh_play_input(s: *mut NhSim, input: u32, series: *const f32, len: u64) -> c_int;
    pub fn nh_loop_input(s: *mut NhSim, input: u32, series: *const f32, len: u64) -> c_int;
    pub fn nh_set_external(s: *mut NhSim, neuron: u32, amps: f32) -> c_int;
    pub fn nh_set_external_all(s: *mut NhSim, ext: *const f32, n: u32) -> c_int;
    pub fn nh_read_external(s: *mut NhSim, out: *mut f32, n: u32) -> c_int;
    pub fn nh_set_noise(s: *mut NhSim, rate_hz: f32, kick_mv: f32) -> c_int;
    pub fn nh_set_noise_sd(s: *mut NhSim, sd: f32) -> c_int;
    pub fn nh_set_v_watch(s: *mut NhSim, ids: *const u32, n_ids: u32) -> c_int;
    pub fn nh_drain_watch(s: *mut NhSim, out: *mut f32, cap_samples: u32, n_samples: *mut u32) -> c_int;
    pub fn nh_set_grid_watches(s: *mut NhSim, n_groups: u32, sizes: *const u32, members: *const u32) -> c_int;
    pub fn nh_grid_traces(s: *mut NhSim, tau_ms: f32, out: *mut f32, cap: u64) -> c_int;
    pub fn nh_read_spikes(
        s: *mut NhSim,
        step0: *mut u64,
        counts: *mut u32,
        max_steps: u64,
        ids: *mut u32,
        max_ids: u64,
        n_steps: *mut u64,
        n_ids: *mut u64,
    ) -> c_int;
    pub fn nh_last_run_ms(s: *mut NhSim, ms: *mut f64) -> c_int;
    pub fn nh_last_run_tail(s: *mut NhSim, ms: *mut f64, launches: *mut u64) -> c_int;
    pub fn nh_profile_steps(
        s: *mut NhSim,
        n_steps: u64,
        main_ms: *mut f64,
        main_n: *mut u64,
        aux_ms: *mut f64,
        aux_n: *mut u64,
    ) -> c_int;
    pub fn nh_autotune(s: *mut NhSim, trial_steps: u32, report: *mut c_char, report_cap: u32) -> c_int;
    pub fn nh_launch_shape(s: *mut NhSim, shape: *mut i32) -> c_int;
    pub fn nh_test_pair_wait(s: *mut NhSim, spin_limit: u32, mute_wave0: u32, fallbacks: *mut u64) -> c_int;
    pub fn nh_probe_copy(bytes: u64, reps: u32, lane_bytes: u32, ms: *mut f64) -> c_int;
    pub fn nh_probe_launch(blocks: u32, launches: u32, us: *mut f64) -> c_int;

    // ---- networks: generators and .ne2 files
    pub fn nh_net_free(net: *mut NhNet);
    pub fn nh_net_view(net: *const NhNet, out: *mut NhNetDesc) -> c_int;
    pub fn nh_net_bench(n: u32, k: u32, seed: u64, w_scale: f32, w_inh: f32, dt_ms: f32, out: *mut *mut NhNet) -> c_int;
    pub fn nh_net_cube(n: u32, mean_deg: f64, seed: u64, w_scale: f32, out: *mut *mut NhNet) -> c_int;
    pub fn nh_net_grid(
        ex: u32,
        ey: u32,
        ez: u32,
        scale: f64,
        lambda: f64,
        inh_frac: f64,
        w_exc: f32,
        w_inh: f32,
        model: c_int,
        n_inputs: u32,
        per_input: u32,
        w_in: f32,
        seed: u64,
        dt_ms: f32,
        out: *mut *mut NhNet,
    ) -> c_int;
    pub fn nh_rng_stream_new(seed: u64, n_src: u32, p: f32, out: *mut *mut NhRngStream) -> c_int;
    pub fn nh_rng_stream_free(s: *mut NhRngStream);
    pub fn nh_rng_stream_window(
        s: *mut NhRngStream,
        steps: u32,
        event: c_int,
        pairs: *const u32,
        n_pairs: u32,
        strength: f32,
        out: *mut f32,
    ) -> c_int;
    pub fn nh_net_load_ne2(bytes: *const u8, len: u64, dt_ms: f32, out: *mut *mut NhNet, info: *mut NhNe2Info) -> c_int;
    pub fn nh_net_type_name(net: *const NhNet, k: u32, buf: *mut c_char, cap: u32) -> c_int;
    pub fn nh_net_neuron_types(net: *const NhNet, ty: *mut u16, n: u32) -> c_int;
    pub fn nh_net_input_name(net: *const NhNet, k: u32, buf: *mut c_char, cap: u32) -> c_int;
    pub fn nh_net_group(
        net: *const NhNet,
        k: u32,
        name: *mut c_char,
        name_cap: u32,
        members: *mut u32,
        cap: u32,
        n_members: *mut u32,
    ) -> c_int;

    // ---- batched form
    pub fn nh_batch_create(n_nets: u32, nets: *const NhNetDesc, cfgs: *const NhConfig, out: *mut *mut NhBatch) -> c_int;
    pub fn nh_batch_destroy(b: *mut NhBatch);
    pub fn nh_batch_run_steps(b: *mut NhBatch, n_steps: u64) -> c_int;
    pub fn nh_batch_sync(b: *mut NhBatch) -> c_int;
    pub fn nh_batch_spikes_total(b: *mut NhBatch, net: u32, out: *mut u64) -> c_int;
    pub fn nh_batch_events_total(b: *mut NhBatch, out: *mut u64) -> c_int;
    pub fn nh_batch_mark_counters(b: *mut NhBatch) -> c_int;
    pub fn nh_batch_counts_since_mark(b: *mut NhBatch, spikes: *mut u64, events: *mut u64) -> c_int;
    pub fn nh_batch_read_state(b: *mut NhBatch, net: u32, slot: u32, out: *mut f32, n: u32) -> c_int;
    pub fn nh_batch_read_spikes(
        b: *mut NhBatch,
        net: u32,
        step0: *mut u64,
        counts: *mut u32,
        max_steps: u64,
        ids: *mut u32,
        max_ids: u64,
        n_steps: *mut u64,
        n_ids: *mut u64,
    ) -> c_int;
    pub fn nh_batch_play_input(b: *mut NhBatch, net: u32, input: u32, series: *const f32, len: u64) -> c_int;
    pub fn nh_batch_spike_counts(b: *mut NhBatch, net: u32, rows: u64, out: *mut u32, n: u32) -> c_int;
    pub fn nh_batch_last_run_ms(b: *mut NhBatch, ms: *mut f64) -> c_int;
    pub fn nh_batch_autotune(b: *mut NhBatch, trial_steps: u32, report: *mut c_char, report_cap: u32) -> c_int;

    // ---- served form
    pub fn nh_server_config_default(c: *mut NhServerConfig);
    pub fn nh_server_create(cfg: *const NhServerConfig, out: *mut *mut NhServer) -> c_int;
    pub fn nh_server_destroy(s: *mut NhServer);
    pub fn nh_server_submit(
        s: *mut NhServer,
        net: *const NhNetDesc,
        cfg: *const NhConfig,
        stim: *const f32,
        stim_steps: u64,
        total_steps: u64,
        win_steps: u32,
        id: *mut u64,
    ) -> c_int;
    pub fn nh_server_wait(
        s: *mut NhServer,
        id: u64,
        timeout_ms: u32,
        counts: *mut u32,
        cap: u64,
        st: *mut NhRequestStats,
    ) -> c_int;
    pub fn nh_server_get_stats(s: *mut NhServer, out: *mut NhServerStats) -> c_int;
}

/// The message of the last failure on this thread.
pub fn last_error() -> String {
    // SAFETY: nh_last_error returns a NUL-terminated thread-local string
    // that stays valid until the next failing call on this thread.
    unsafe {
        let p = nh_last_error();
        if p.is_null() {
            String::new()
        } else {
            std::ffi::CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}
