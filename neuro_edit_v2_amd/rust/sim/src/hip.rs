//! gfx950 backend of the `sim` crate: `HipSim` implements `Backend`
//! (sim/src/lib.rs:119-184) over libneurohip's C ABI (`sim-hip-sys`), in
//! place of `CudaSim` (sim/src/cuda.rs).  The file goes next to cpu.rs as
//! `sim/src/hip.rs` behind a `hip` feature; `Sim::with_params`
//! (lib.rs:217-233) builds it for `BackendKind::Cuda`, so every selector the
//! reference already has (`--cuda`, `--int-weights`, `proto::Backend::Cuda`)
//! reaches it unchanged.
//!
//! Construction mirrors `CudaSim::with_params` (cuda.rs:130-239): the
//! parameter check, `synapse_csr` at the config's dt, the stimulus map; the
//! flat arrays are handed over as pointers and uploaded once.  The stimulus
//! program and the external currents stay host-side in the library as in
//! cuda.rs:241-252; this struct keeps mirrors of both for `stimulus()` and
//! `external()`, which return owned copies like every getter of the trait.
//! Readback and synchronisation failures panic, as in CudaSim
//! (cuda.rs:612-743 unwrap/expect), where the trait has no error channel.
use net_format::Network;
use sim_hip_sys as ffi;

use crate::state::{NeuronParams, SimConfig};
use crate::stim::{StimMap, Stimulus};
use crate::{Backend, BackendKind};

pub struct HipSim {
    h: *mut ffi::NhSim,
    n: usize,
    n_state: usize,
    n_inputs: usize,
    n_watch: usize,
    grid_sizes: Vec<usize>,
    stim: Stimulus,
    ext: Vec<f32>,
    cfg: SimConfig,
}

fn status(rc: i32) -> Result<(), String> {
    if rc == ffi::NH_OK {
        Ok(())
    } else {
        Err(format!("hip backend: {} (status {rc})", ffi::last_error()))
    }
}

fn check(rc: i32) -> anyhow::Result<()> {
    status(rc).map_err(anyhow::Error::msg)
}

/// True when libneurohip finds a usable device (`cuda_available`,
/// lib.rs:192-203).
pub fn hip_available() -> bool {
    // SAFETY: a device probe without arguments.
    unsafe { ffi::nh_available() != 0 }
}

impl HipSim {
    pub fn new(net: &Network, cfg: SimConfig) -> anyhow::Result<Self> {
        let params = NeuronParams::from_network(net).map_err(anyhow::Error::msg)?;
        Self::with_params(net, cfg, params)
    }

    pub fn with_params(net: &Network, cfg: SimConfig, params: NeuronParams) -> anyhow::Result<Self> {
        let n = net.neurons.len();
        params.check(n).map_err(anyhow::Error::msg)?;
        let csr = net.synapse_csr(cfg.dt_ms).map_err(anyhow::Error::msg)?;
        let sm = StimMap::from_network(net);
        // registry id per entry of the net's model table (model/mod.rs:118-125);
        // the LIF test model (sim/tests/lif.rs:14-104), which is not in the
        // registry, is built into this backend as id 2
        let mut ids = Vec::with_capacity(params.models.len());
        for m in &params.models {
            let id = match crate::model::by_id(m.id()) {
                Some((id, _)) => id,
                None if m.id() == "lif" => 2,
                None => anyhow::bail!("model {} is not built into this backend", m.id()),
            };
            ids.push(id);
        }
        // state rows are two wide on the device (a LIF-only net has one slot)
        let x0: std::borrow::Cow<[f32]> = if params.n_state >= 2 {
            std::borrow::Cow::Borrowed(&params.x0)
        } else {
            std::borrow::Cow::Owned(params.x0.iter().flat_map(|&v| [v, 0.0]).collect())
        };
        let any_tau = csr.tau_ms.iter().any(|&t| t > 0.0);
        let desc = ffi::NhNetDesc {
            n: n as u32,
            n_syn: csr.tgt.len() as u64,
            src_ptr: csr.src_ptr.as_ptr(),
            tgt: csr.tgt.as_ptr(),
            w: csr.w.as_ptr(),
            delay: csr.delay_steps.as_ptr(),
            tau: if any_tau { csr.tau_ms.as_ptr() } else { std::ptr::null() },
            n_models: ids.len() as u32,
            model_ids: ids.as_ptr(),
            model: params.model.as_ptr(),
            n_params: params.n_params as u32,
            n_state: params.n_state.max(2) as u32,
            p: params.p.as_ptr(),
            x0: x0.as_ptr(),
            n_inputs: net.inputs.len() as u32,
            stim_ptr: sm.neuron_ptr.as_ptr(),
            stim_input: sm.input_idx.as_ptr(),
            stim_w: sm.weight.as_ptr(),
        };
        let mut c = std::mem::MaybeUninit::<ffi::NhConfig>::uninit();
        // SAFETY: nh_config_default writes every field.
        let mut c = unsafe {
            ffi::nh_config_default(c.as_mut_ptr());
            c.assume_init()
        };
        c.dt_ms = cfg.dt_ms;
        c.synapse_tau_ms = cfg.synapse_tau_ms;
        c.readout_tau_ms = cfg.readout_tau_ms;
        c.noise_rate_hz = cfg.noise_rate_hz;
        c.noise_kick_mv = cfg.noise_kick_mv;
        c.seed = cfg.seed;
        c.record_spikes = cfg.record_spikes as i32;
        c.spike_record_steps = cfg.spike_record_steps as u32;
        c.int_weights = cfg.int_weights as i32;
        c.int_scale = 4096.0; // WSCALE, kernels.rs:54-66
        let mut h = std::ptr::null_mut();
        // SAFETY: every pointer in desc refers to a live Vec of the stated
        // length for the duration of the call; nh_create copies what it keeps.
        check(unsafe { ffi::nh_create(&desc, &c, &mut h) })?;
        Ok(HipSim {
            h,
            n,
            n_state: params.n_state,
            n_inputs: net.inputs.len(),
            n_watch: 0,
            grid_sizes: Vec::new(),
            stim: Stimulus::for_network(net),
            ext: vec![0.0; n],
            cfg,
        })
    }

    /// Gaussian current sd per step (an addition to the reference's kick noise).
    pub fn set_noise_sd(&mut self, sd: f32) -> anyhow::Result<()> {
        check(unsafe { ffi::nh_set_noise_sd(self.h, sd) })
    }

    /// The full raster since the last drain: (first step, spikes per step,
    /// ascending ids per step) -- an addition over the trait.
    pub fn read_spikes(&mut self) -> anyhow::Result<(u64, Vec<u32>, Vec<u32>)> {
        let (mut s0, mut ns, mut ni) = (0u64, 0u64, 0u64);
        let rc = unsafe {
            ffi::nh_read_spikes(self.h, &mut s0, std::ptr::null_mut(), 0, std::ptr::null_mut(), 0, &mut ns, &mut ni)
        };
        if rc != ffi::NH_OK && rc != ffi::NH_ERR_CAPACITY {
            check(rc)?;
        }
        let mut counts = vec![0u32; ns as usize];
        let mut ids = vec![0u32; ni as usize];
        check(unsafe {
            ffi::nh_read_spikes(self.h, &mut s0, counts.as_mut_ptr(), ns, ids.as_mut_ptr(), ni, &mut ns, &mut ni)
        })?;
        Ok((s0, counts, ids))
    }
}

impl Drop for HipSim {
    fn drop(&mut self) {
        // SAFETY: h came from nh_create and is destroyed once.
        unsafe { ffi::nh_destroy(self.h) }
    }
}

impl Backend for HipSim {
    fn kind(&self) -> BackendKind {
        BackendKind::Cuda
    }
    fn n(&self) -> usize {
        self.n
    }
    fn step_count(&self) -> u64 {
        unsafe { ffi::nh_step_count(self.h) }
    }
    fn time_ms(&self) -> f64 {
        unsafe { ffi::nh_time_ms(self.h) }
    }
    fn run_steps(&mut self, n: usize) -> anyhow::Result<()> {
        check(unsafe { ffi::nh_run_steps(self.h, n as u64) })
    }
    fn run_ms(&mut self, ms: f32) -> anyhow::Result<()> {
        check(unsafe { ffi::nh_run_ms(self.h, ms) })
    }
    fn reset_state(&mut self) {
        check(unsafe { ffi::nh_reset(self.h) }).expect("hip reset");
    }
    fn synchronize(&self) {
        check(unsafe { ffi::nh_sync(self.h) }).expect("hip synchronize");
    }
    fn dev_mem_bytes(&self) -> usize {
        unsafe { ffi::nh_dev_mem(self.h) as usize }
    }
    fn trace(&self) -> Vec<f32> {
        let mut v = vec![0.0f32; self.n];
        check(unsafe { ffi::nh_read_trace(self.h, v.as_mut_ptr(), self.n as u32) }).expect("trace");
        v
    }
    fn state(&self, slot: usize) -> Vec<f32> {
        if slot >= self.n_state || self.n == 0 {
            return Vec::new();
        }
        let mut v = vec![0.0f32; self.n];
        check(unsafe { ffi::nh_read_state(self.h, slot as u32, v.as_mut_ptr(), self.n as u32) }).expect("state");
        v
    }
    fn last_spikes(&self) -> Vec<u32> {
        let mut ids = vec![0u32; self.n];
        let mut k = 0u32;
        check(unsafe { ffi::nh_last_spikes(self.h, ids.as_mut_ptr(), self.n as u32, &mut k) }).expect("last_spikes");
        ids.truncate(k as usize);
        ids
    }
    fn spike_counts(&self, rows: usize) -> Result<Vec<u32>, String> {
        let mut v = vec![0u32; self.n];
        let rc = unsafe { ffi::nh_spike_counts(self.h, rows as u64, v.as_mut_ptr(), self.n as u32) };
        if rc == ffi::NH_ERR_WINDOW {
            // the reference's own message (cpu.rs:463-468)
            return Err(ffi::last_error());
        }
        status(rc).map(|_| v)
    }
    fn spike_row(&self, back: usize) -> Option<Vec<u64>> {
        let mut w = vec![0u64; self.n.div_ceil(64)];
        match unsafe { ffi::nh_spike_row(self.h, back as u64, w.as_mut_ptr(), w.len() as u32) } {
            ffi::NH_OK => Some(w),
            ffi::NH_ERR_NONE => None,
            rc => panic!("{}", status(rc).unwrap_err()),
        }
    }
    fn spikes_total(&self) -> u64 {
        let mut v = 0u64;
        check(unsafe { ffi::nh_spikes_total(self.h, &mut v) }).expect("spikes_total");
        v
    }
    fn mean_rate_hz(&self) -> f64 {
        let mut v = 0f64;
        check(unsafe { ffi::nh_mean_rate_hz(self.h, &mut v) }).expect("mean_rate_hz");
        v
    }
    fn set_input_constant(&mut self, input: usize, value: f32) {
        if input < self.n_inputs {
            self.stim.constants[input] = value;
            check(unsafe { ffi::nh_set_input_constant(self.h, input as u32, value) }).expect("set_input_constant");
        }
    }
    fn input_values(&self) -> Vec<f32> {
        let mut v = vec![0.0f32; self.n_inputs];
        check(unsafe { ffi::nh_input_values(self.h, v.as_mut_ptr(), self.n_inputs as u32) }).expect("input_values");
        v
    }
    fn stimulus(&self) -> Stimulus {
        self.stim.clone()
    }
    fn external(&self) -> Vec<f32> {
        self.ext.clone()
    }
    fn play_input(&mut self, input: usize, series: Vec<f32>) {
        if input >= self.n_inputs {
            return;
        }
        check(unsafe { ffi::nh_play_input(self.h, input as u32, series.as_ptr(), series.len() as u64) })
            .expect("play_input");
        self.stim.play(input, self.step_count(), series);
    }
    fn loop_input(&mut self, input: usize, series: Vec<f32>) {
        if input >= self.n_inputs {
            return;
        }
        check(unsafe { ffi::nh_loop_input(self.h, input as u32, series.as_ptr(), series.len() as u64) })
            .expect("loop_input");
        self.stim.play_looped(input, self.step_count(), series);
    }
    fn set_v_watch(&mut self, ids: &[u32]) {
        check(unsafe { ffi::nh_set_v_watch(self.h, ids.as_ptr(), ids.len() as u32) }).expect("set_v_watch");
        self.n_watch = ids.len();
    }
    fn drain_watch(&mut self) -> anyhow::Result<Vec<Vec<Vec<f32>>>> {
        // [state slot][watch][samples], at most 4096 samples kept per watch
        const CAP: usize = 4096;
        let (nx, nw) = (self.n_state, self.n_watch);
        // the device keeps two state slots per watch (a LIF-only net reads one)
        let mut out = vec![0.0f32; nx.max(2) * nw * CAP];
        let mut k = 0u32;
        check(unsafe { ffi::nh_drain_watch(self.h, out.as_mut_ptr(), CAP as u32, &mut k) })?;
        let k = k as usize;
        Ok((0..nx)
            .map(|s| (0..nw).map(|w| out[(s * nw + w) * CAP..(s * nw + w) * CAP + k].to_vec()).collect())
            .collect())
    }
    fn set_grid_watches(&mut self, groups: &[Vec<u32>]) {
        let sizes: Vec<u32> = groups.iter().map(|g| g.len() as u32).collect();
        let members: Vec<u32> = groups.iter().flatten().copied().collect();
        check(unsafe { ffi::nh_set_grid_watches(self.h, sizes.len() as u32, sizes.as_ptr(), members.as_ptr()) })
            .expect("set_grid_watches");
        self.grid_sizes = groups.iter().map(|g| g.len()).collect();
    }
    fn grid_traces(&mut self, tau_ms: f32) -> anyhow::Result<Vec<Vec<f32>>> {
        let total: usize = self.grid_sizes.iter().sum();
        let mut flat = vec![0.0f32; total];
        check(unsafe { ffi::nh_grid_traces(self.h, tau_ms, flat.as_mut_ptr(), total as u64) })?;
        let mut out = Vec::with_capacity(self.grid_sizes.len());
        let mut off = 0;
        for &m in &self.grid_sizes {
            out.push(flat[off..off + m].to_vec());
            off += m;
        }
        Ok(out)
    }
    fn set_external(&mut self, neuron: usize, amps: f32) {
        if neuron < self.n {
            self.ext[neuron] = amps;
            check(unsafe { ffi::nh_set_external(self.h, neuron as u32, amps) }).expect("set_external");
        }
    }
    fn set_external_all(&mut self, ext: &[f32]) {
        let m = ext.len().min(self.n);
        self.ext[..m].copy_from_slice(&ext[..m]);
        check(unsafe { ffi::nh_set_external_all(self.h, self.ext.as_ptr(), self.n as u32) }).expect("set_external_all");
    }
    fn set_noise(&mut self, rate_hz: f32, kick_mv: f32) {
        self.cfg.noise_rate_hz = rate_hz;
        self.cfg.noise_kick_mv = kick_mv;
        check(unsafe { ffi::nh_set_noise(self.h, rate_hz, kick_mv) }).expect("set_noise");
    }
}

// A HipSim is used from one thread at a time, like CudaSim (script.rs:21
// holds it in an Rc<RefCell<Sim>>); the library handle may move between
// threads (the server's job threads, jobs.rs:156-159).
unsafe impl Send for HipSim {}
