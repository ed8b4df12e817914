/*
 * neuro_hip.h -- C ABI of the MI355X (gfx950) simulation backend.
 *
 * This is the drop-in boundary for the reference's `sim` crate backend
 * trait (sim/src/lib.rs:119-184) and its construction path
 * (sim/src/lib.rs:205-233).  Every function takes plain pointers and sizes;
 * outputs are caller-allocated with explicit lengths; no exception or panic
 * crosses the ABI.  Each function returns an nh_status (0 = ok); the text of
 * the last failure on the calling thread is nh_last_error().
 *
 * Threading: one nh_sim per thread (or serialise calls); each nh_sim owns
 * one HIP stream.  nh_run_steps is asynchronous; readers synchronise.
 *
 * Citations name the reference item each entry point replaces.
 */
#ifndef NEURO_HIP_H
#define NEURO_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NH_OK = 0,
    NH_ERR_ARG = 1,          /* bad argument / size */
    NH_ERR_HIP = 2,          /* HIP runtime failure */
    NH_ERR_NO_DEVICE = 3,    /* no usable GPU */
    NH_ERR_WINDOW = 4,       /* spike_counts window between ring depth and t */
    NH_ERR_NONE = 5,         /* Option::None (e.g. spike_row outside window) */
    NH_ERR_UNSUPPORTED = 6,  /* feature not provided by this backend */
    NH_ERR_CAPACITY = 7,     /* admission / buffer capacity exceeded */
    NH_ERR_NET = 8           /* invalid network (e.g. delay past u16 ring) */
} nh_status;

/* Flat network layout the simulators run: the reference's SynCsr
 * (net-format/src/network.rs:431-441, built by synapse_csr :654-692),
 * NeuronParams (sim/src/state.rs:57-74) and StimMap (sim/src/stim.rs:26-31). */
typedef struct {
    uint32_t n;
    uint64_t n_syn;
    const uint32_t *src_ptr;   /* n+1, outgoing synapses of i: src_ptr[i]..src_ptr[i+1] */
    const uint32_t *tgt;       /* n_syn */
    const float *w;            /* n_syn, signed by Dale's law */
    const uint16_t *delay;     /* n_syn, steps >= 1 */
    const float *tau;          /* n_syn or NULL; > 0 = custom-tau synapse */
    uint32_t n_models;         /* model table size (<= 8) */
    const uint8_t *model_ids;  /* registry id per table entry: 0 izh, 1 adex, 2 lif (the
                                  reference's test-only plug-in, sim/tests/lif.rs:14-104:
                                  row [tau, e, r, vth, vreset], state [v, unused]) */
    const uint8_t *model;      /* n, index into the model table */
    uint32_t n_params;         /* param row width (max over the table) */
    uint32_t n_state;          /* state row width (2 for izh and adex) */
    const float *p;            /* n * n_params */
    const float *x0;           /* n * n_state */
    uint32_t n_inputs;
    const uint32_t *stim_ptr;  /* n+1 or NULL */
    const uint32_t *stim_input;
    const float *stim_w;
} nh_net_desc;

/* SimConfig (sim/src/state.rs:7-49) plus this backend's additions. */
typedef struct {
    float dt_ms;
    float synapse_tau_ms;
    float readout_tau_ms;
    float noise_rate_hz;       /* Poisson kicks (reference `--noise`) */
    float noise_kick_mv;
    float noise_sd;            /* Gaussian current sd per step (addition) */
    uint64_t seed;
    int32_t record_spikes;
    uint32_t spike_record_steps;
    int32_t int_weights;       /* fixed-point delay ring (sim/src/kernels.rs:57-66) */
    float int_scale;           /* 4096 = reference, 65536 = Q16.16 */
    uint32_t raster_steps;     /* > 0: keep the full spike raster; the device holds a
                                  window of raster_steps steps (one bit per neuron per
                                  step), drained into host memory when it fills and on
                                  nh_read_spikes */
    int32_t exec_mode;         /* 0 auto, 1 direct launches, 2 hipGraph chunks */
    int32_t step_shape;        /* -1 auto, 0 two kernels per step (neuron, scatter),
                                  1 one fused kernel per step (update + own-spike push),
                                  2 two steps per launch (the delay-1 arrivals of the
                                  first step pulled by their targets' waves); used where
                                  the whole grid fits on the device at once and no
                                  neuron is watched and no synapse has its own tau,
                                  otherwise runs as 1.  3 many steps per launch (one
                                  persistent kernel per chunk of nh_run_steps; the
                                  delay-1 arrivals pulled, the rest gated on the
                                  launch's own progress tags): one network without
                                  stimulus inputs, watches or custom-tau synapses
                                  whose grid fits on the device at once, otherwise
                                  runs as 1; same results, never chosen by auto
                                  (slower per step than 2 at every run length
                                  measured).  Auto = 2 where at most 256
                                  delay-1 synapses end in any 64-neuron group, else 1 */
    uint32_t graph_steps;      /* steps per captured graph (0 = 64) */
    uint32_t block;            /* workgroup size of the step kernels, 64..512 (0 = 256) */
    uint32_t lds_bins;         /* one-step shape (1): push through a workgroup table of
                                  this many LDS bins (a power of two, 256..8192), one
                                  HBM atomic per distinct ring word; 0 = each spiking
                                  wave pushes its own rows */
} nh_config;

/* SimConfig::default() values plus neutral additions. */
void nh_config_default(nh_config *cfg);

const char *nh_last_error(void);

/* cuda_available -- sim/src/lib.rs:192-203 */
int nh_available(void);

typedef struct nh_sim nh_sim;

/* Sim::with_params(net, cfg, BackendKind::Cuda, params) -- sim/src/lib.rs:217-233,
 * CudaSim::with_params -- sim/src/cuda.rs:130-239 */
/* Host-only checks of a network and config, made by nh_create (and the
 * batched/served entry points) before any device call: dt range
 * (sim/src/lib.rs:35-39), the row-table consistency of NeuronParams::check
 * (sim/src/state.rs:197-217), and the per-synapse / per-input checks of
 * Network::validate (net-format/src/network.rs:1074-1150) on the flat
 * arrays (targets in range, delays >= 1, finite weights and taus, int ring
 * range).  NH_OK or NH_ERR_ARG with the reason in nh_last_error. */
int nh_check_net(const nh_net_desc *net, const nh_config *cfg);
int nh_create(const nh_net_desc *net, const nh_config *cfg, nh_sim **out);
void nh_destroy(nh_sim *s);

/* Backend::run_steps / run_ms -- sim/src/lib.rs:126-127, cuda.rs:584-609 */
int nh_run_steps(nh_sim *s, uint64_t n_steps);
int nh_run_ms(nh_sim *s, float ms);
/* Backend::reset_state -- lib.rs:129, cuda.rs:615-644 */
int nh_reset(nh_sim *s);
/* Backend::synchronize -- lib.rs:131 */
int nh_sync(nh_sim *s);
/* Backend::n / step_count / time_ms -- lib.rs:123-125 */
uint32_t nh_n(const nh_sim *s);
uint64_t nh_step_count(const nh_sim *s);
double nh_time_ms(const nh_sim *s);
/* Backend::dev_mem_bytes -- lib.rs:133-135, cuda.rs:804-820 */
uint64_t nh_dev_mem(const nh_sim *s);

/* Backend::trace -- lib.rs:137 (out: n floats) */
int nh_read_trace(nh_sim *s, float *out, uint32_t n);
/* Backend::state(slot) -- lib.rs:139 (out: n floats) */
int nh_read_state(nh_sim *s, uint32_t slot, float *out, uint32_t n);
/* Decaying synaptic current s[i] (CpuSim::s, cpu.rs:33); parity probes. */
int nh_read_syn_current(nh_sim *s, float *out, uint32_t n);
/* Backend::spikes_total / mean_rate_hz -- lib.rs:149-151 */
int nh_spikes_total(nh_sim *s, uint64_t *out);
int nh_mean_rate_hz(nh_sim *s, double *out);
/* Per-neuron lifetime spike totals (CpuSim::pn_total, cpu.rs:60). */
int nh_read_pn_total(nh_sim *s, uint64_t *out, uint32_t n);
/* Synaptic events delivered so far (one per spike x out-synapse). */
int nh_events_total(nh_sim *s, uint64_t *out);
/* Measurement helpers (additions, no counterpart in the reference): a
 * snapshot of the per-neuron spike totals taken on the device in stream
 * order -- the host does not wait, so a timed run enqueued right after it
 * starts on a busy device -- and the spikes and synaptic events since the
 * last snapshot (synchronises).  A reset clears the snapshot
 * (NH_ERR_ARG until the next mark). */
int nh_mark_counters(nh_sim *s);
int nh_counts_since_mark(nh_sim *s, uint64_t *spikes, uint64_t *events);
/* Backend::spike_row(back) -- lib.rs:148; NH_ERR_NONE for None. */
int nh_spike_row(nh_sim *s, uint64_t back, uint64_t *words, uint32_t n_words);
/* Backend::last_spikes -- lib.rs:142 (ascending ids) */
int nh_last_spikes(nh_sim *s, uint32_t *ids, uint32_t cap, uint32_t *n_out);
/* Backend::spike_counts(rows) -- lib.rs:146; NH_ERR_WINDOW for the
 * "exceeds the recorded steps" error (cpu.rs:463-468).  Replaces
 * CudaSim::spike_counts (cuda.rs:725-763, which copies the whole record
 * ring): the window is counted on the device (a count buffer of n u32 is
 * allocated at the first windowed call) and only n counts are copied. */
int nh_spike_counts(nh_sim *s, uint64_t rows, uint32_t *out, uint32_t n);
/* nh_spike_counts that allocates no device memory -- no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The window is
 * counted into a 16 MiB library scratch per GPU (at most 4,194,304 neurons;
 * larger networks get NH_ERR_CAPACITY).  Calls that overlap on one GPU race
 * on that scratch. */
int nh_spike_counts_noalloc(nh_sim *s, uint64_t rows, uint32_t *out, uint32_t n);

/* Backend::set_input_constant / input_values / play_input / loop_input --
 * lib.rs:152-163 (stimulus program kept host-side as in cuda.rs:241-252) */
int nh_set_input_constant(nh_sim *s, uint32_t input, float value);
int nh_input_values(nh_sim *s, float *out, uint32_t n_inputs);
int nh_play_input(nh_sim *s, uint32_t input, const float *series, uint64_t len);
int nh_loop_input(nh_sim *s, uint32_t input, const float *series, uint64_t len);
/* Backend::set_external / set_external_all / external -- lib.rs:158-181 */
int nh_set_external(nh_sim *s, uint32_t neuron, float amps);
int nh_set_external_all(nh_sim *s, const float *ext, uint32_t n);
int nh_read_external(nh_sim *s, float *out, uint32_t n);
/* Backend::set_noise -- lib.rs:183 */
int nh_set_noise(nh_sim *s, float rate_hz, float kick_mv);
/* Gaussian-current sd on a live sim (addition). */
int nh_set_noise_sd(nh_sim *s, float sd);

/* Backend::set_v_watch / drain_watch -- lib.rs:166-170, cuda.rs:531-570.
 * drain: out is [state slot][watch][cap_samples]; *n_samples gets the count
 * per watch (oldest first, at most 4096 kept). */
int nh_set_v_watch(nh_sim *s, const uint32_t *ids, uint32_t n_ids);
int nh_drain_watch(nh_sim *s, float *out, uint32_t cap_samples, uint32_t *n_samples);

/* Backend::set_grid_watches / grid_traces -- lib.rs:173-176.  members is
 * the concatenation of the groups; out is the concatenation of the
 * per-texel traces. */
int nh_set_grid_watches(nh_sim *s, uint32_t n_groups, const uint32_t *sizes,
                        const uint32_t *members);
int nh_grid_traces(nh_sim *s, float tau_ms, float *out, uint64_t cap);

/* Full spike raster (addition; the reference keeps only a rolling bit
 * ring).  Drains the host-side raster accumulated since the last call:
 * counts[k] = spikes at step step0+k, ids = concatenated ids per step in
 * ascending order.  Returns NH_ERR_CAPACITY (and drains nothing) if the
 * caller's arrays are too small; *n_steps / *n_ids always get the sizes. */
int nh_read_spikes(nh_sim *s, uint64_t *step0, uint32_t *counts, uint64_t max_steps,
                   uint32_t *ids, uint64_t max_ids, uint64_t *n_steps, uint64_t *n_ids);

/* Device timing of the step kernels over the last nh_run_steps call
 * (hipEvents on the sim's stream): total milliseconds. */
int nh_last_run_ms(nh_sim *s, double *ms);
/* The last run's time after its first step launch (addition; no
 * counterpart in the reference): milliseconds from the end of its first
 * step launch (or first graph replay) to the end of the run (which holds
 * any one-thread step-counter advance launches as well: chunks with
 * stimulus inputs and graph replays have them), and the step launches in
 * that span (a paired launch counts once).  This is the
 * back-to-back time per launch without the host's latency before the first
 * kernel.  *launches is 0 (and *ms the whole run's time) when the run had
 * no launch after its first. */
int nh_last_run_tail(nh_sim *s, double *ms, uint64_t *launches);
/* Run n_steps with direct launches and a hipEvent pair around every
 * launch of the step kernels (on the sim's stream).  main_ms / main_n: the
 * neuron (or fused step) kernel; aux_ms / aux_n: the scatter kernel
 * (0 in the fused shape).  Same results as nh_run_steps. */
int nh_profile_steps(nh_sim *s, uint64_t n_steps, double *main_ms, uint64_t *main_n,
                     double *aux_ms, uint64_t *aux_n);

/* Autotuner (addition): from step 0, time every candidate launch shape
 * (one fused kernel or two kernels per step, or two steps per launch where
 * that shape can run) x workgroup size (64..512) x captured-graph length
 * over trial_steps, keep the fastest and reset the sim.  A text report of
 * the trials goes to report (may be NULL). */
int nh_autotune(nh_sim *s, uint32_t trial_steps, char *report, uint32_t report_cap);
/* Launch shape the next nh_run_steps uses (addition): 0 two kernels per
 * step, 1 one fused kernel per step, 2 two steps per launch, 3 many steps
 * per launch. */
int nh_launch_shape(nh_sim *s, int32_t *shape);
/* Test hook for the give-up paths of the shapes that wait inside a launch
 * (no counterpart in the reference): spin_limit = polls before a waiting
 * wave gives up (0 = the default).  mute_wave0 = 1: in the paired shape wave
 * 0 starts each launch 1 ms late, as a wave that is not resident would, so
 * the waves it feeds give up waiting and recompute its first step
 * themselves, and the results are unchanged; in the persistent shape wave 0
 * withholds its published spike words.  Persistent shape: the launch ends early and the
 * next synchronising call reports NH_ERR_HIP (nh_reset clears it).  The
 * hook synchronises first; *fallbacks (may be null) gets the number of
 * synchronisations so far that found a paired launch's wait given up.
 * Never set outside tests. */
int nh_test_pair_wait(nh_sim *s, uint32_t spin_limit, uint32_t mute_wave0, uint64_t *fallbacks);

/* Measurement helpers (additions, no counterpart in the reference).
 * nh_probe_copy: copy `bytes` between two device buffers with a plain
 * load/store kernel of `lane_bytes` (4 or 16) per lane, `reps` times; *ms
 * gets the average kernel time.  A copy of known size reads and writes
 * exactly `bytes`, which calibrates the HBM byte counters.
 * nh_probe_launch: replay a graph of `launches` empty kernels of `blocks`
 * x 256 threads; *us gets the time per kernel (the one-launch-per-step
 * floor). */
int nh_probe_copy(uint64_t bytes, uint32_t reps, uint32_t lane_bytes, double *ms);
int nh_probe_launch(uint32_t blocks, uint32_t launches, double *us);

/* ---------------- network generators (net-format restated) ----------------
 * An nh_net owns the flat arrays; nh_net_view fills an nh_net_desc that
 * points into it. */
typedef struct nh_net nh_net;
void nh_net_free(nh_net *net);
int nh_net_view(const nh_net *net, nh_net_desc *out);
/* bench::bench_net -- sim/src/bench.rs:10-53 (fixed out-degree k,
 * constant weights, delays U[1,21) ms exc / 1 ms inh) built at dt. */
int nh_net_bench(uint32_t n, uint32_t k, uint64_t seed, float w_scale, float w_inh,
                 float dt_ms, nh_net **out);
/* Random-cube network: positions U[-500,500)^3, 80% RS / 20% FS,
 * Poisson(mean_deg) out-degree by inversion, uniform non-self targets,
 * w = U[0,1)*w_scale (Dale-signed), delay = clamp(1 + floor(19 d / dmax),1,20)
 * steps with dmax = 1000*sqrt(3). */
int nh_net_cube(uint32_t n, double mean_deg, uint64_t seed, float w_scale, nh_net **out);
/* 3D-grid liquid (builders.rs:88-181 semantics with a spatially bucketed
 * pair loop): edge^3 neurons at spacing 1, p = scale*exp(-(d/lambda)^2),
 * inh_frac Dale split (set_exc_inh), weight w (izh) -- model 0 izh, 1 adex
 * (exc TONIC, inh NONADAP).  n_inputs inputs with `per_input` private
 * uniformly drawn targets of weight w_in. */
int nh_net_grid(uint32_t ex, uint32_t ey, uint32_t ez, double scale, double lambda,
                double inh_frac, float w_exc, float w_inh, int model, uint32_t n_inputs,
                uint32_t per_input, float w_in, uint64_t seed, float dt_ms, nh_net **out);

/* RngStream::window -- sim/src/script.rs:2849-2894: `steps` x `n_src`
 * Bernoulli(p) bits from a seeded Rng; with event != 0 each listed pair
 * (pairs[2k], pairs[2k+1]) shares one Bernoulli(p) bit with probability
 * `strength` per step.  The generator state carries across calls on the
 * same handle (nh_rng_stream_new/free). */
typedef struct nh_rng_stream nh_rng_stream;
int nh_rng_stream_new(uint64_t seed, uint32_t n_src, float p, nh_rng_stream **out);
void nh_rng_stream_free(nh_rng_stream *s);
int nh_rng_stream_window(nh_rng_stream *s, uint32_t steps, int event, const uint32_t *pairs,
                         uint32_t n_pairs, float strength, float *out);

/* .ne2 network files (net-format/src/io.rs:437-468 `io::load`): magic,
 * version 1..7, bincode `Network`, older layouts lifted as the reference's
 * loaders do (io.rs:14-330), validated (network.rs:1007-1150), then turned
 * into the simulator layout with NeuronParams::from_network
 * (sim/src/state.rs:98-153), synapse_csr(dt) (network.rs:654-692) and
 * StimMap::from_network (sim/src/stim.rs:34-57).  dt_ms <= 0 takes the
 * file's meta.dt_ms.  The caller reads the file; `info` (optional) gets the
 * file's meta and table sizes. */
typedef struct {
    uint32_t version;
    float dt_ms, conduction_speed, synapse_tau_ms, readout_tau_ms, noise_rate_hz, noise_kick_mv;
    uint64_t seed;
    uint32_t n_types, n_groups, n_inputs, n_samples, n_readouts, n_scripts;
} nh_ne2_info;
int nh_net_load_ne2(const uint8_t *bytes, uint64_t len, float dt_ms, nh_net **out, nh_ne2_info *info);
/* Tables of a loaded file that callers address by index: type names (the
 * lifted type table for v1-v3 files, io.rs:98-170), each neuron's type,
 * input names (stimulus input k), groups (grid watches take members). */
int nh_net_type_name(const nh_net *net, uint32_t k, char *buf, uint32_t cap);
int nh_net_neuron_types(const nh_net *net, uint16_t *ty, uint32_t n);
int nh_net_input_name(const nh_net *net, uint32_t k, char *buf, uint32_t cap);
int nh_net_group(const nh_net *net, uint32_t k, char *name, uint32_t name_cap, uint32_t *members,
                 uint32_t cap, uint32_t *n_members);

/* ---------------- batched form ----------------
 * Many independent networks packed by offset tables into one launch
 * sequence; each network keeps its own config (dt must match), seed, t
 * origin and stimulus.  Results are per network and equal the plain
 * path's (int mode: bit-exact). */
typedef struct nh_batch nh_batch;
int nh_batch_create(uint32_t n_nets, const nh_net_desc *nets, const nh_config *cfgs,
                    nh_batch **out);
void nh_batch_destroy(nh_batch *b);
int nh_batch_run_steps(nh_batch *b, uint64_t n_steps);
int nh_batch_sync(nh_batch *b);
int nh_batch_spikes_total(nh_batch *b, uint32_t net, uint64_t *out);
int nh_batch_events_total(nh_batch *b, uint64_t *out);
/* nh_mark_counters / nh_counts_since_mark over the whole batch (additions). */
int nh_batch_mark_counters(nh_batch *b);
int nh_batch_counts_since_mark(nh_batch *b, uint64_t *spikes, uint64_t *events);
int nh_batch_read_state(nh_batch *b, uint32_t net, uint32_t slot, float *out, uint32_t n);
int nh_batch_read_spikes(nh_batch *b, uint32_t net, uint64_t *step0, uint32_t *counts,
                         uint64_t max_steps, uint32_t *ids, uint64_t max_ids,
                         uint64_t *n_steps, uint64_t *n_ids);
int nh_batch_play_input(nh_batch *b, uint32_t net, uint32_t input, const float *series,
                        uint64_t len);
/* nh_spike_counts for one packed network (Backend::spike_counts, lib.rs:146;
 * cuda.rs:725-763): record_spikes must be set in the first network's config
 * (the record ring is arena-wide); local neuron ids. */
int nh_batch_spike_counts(nh_batch *b, uint32_t net, uint64_t rows, uint32_t *out, uint32_t n);
int nh_batch_last_run_ms(nh_batch *b, double *ms);
/* nh_autotune for a batch (from step 0; resets the batch). */
int nh_batch_autotune(nh_batch *b, uint32_t trial_steps, char *report, uint32_t report_cap);

/* ---------------- served form ----------------
 * A request loop (a worker thread) over one resident arena: requests are
 * admitted while neurons, synapses and slots are free (capacity fixed at
 * creation from free HBM when left 0), keep their state and delay ring
 * resident, and every step all in-flight requests advance together in the
 * same launch.  A request runs exactly `total_steps` (1..2^32-1) steps from
 * its admission (the kernel stops it at its end even mid-round) and returns
 * per-neuron spike counts for `n_windows` windows of `win_steps`.
 * The reference runs one thread and one simulator per job with no GPU
 * sharing (server/src/jobs.rs:106-161,226-339); this is its batched
 * counterpart. */
typedef struct {
    uint64_t cap_neurons;      /* 0 = from free HBM */
    uint64_t cap_synapses;     /* 0 = from free HBM */
    uint64_t cap_stim;         /* stimulus entries (0 = 64 per neuron slot) */
    uint32_t max_requests;     /* concurrent slots (<= 65535) */
    uint32_t max_delay;        /* ring depth - 1 (<= 31) */
    uint32_t n_inputs;         /* stimulus inputs per request */
    uint32_t n_windows;        /* windowed spike counters per request */
    uint32_t chunk_steps;      /* steps per scheduling round (rounded to the ring depth) */
    int32_t int_weights;
    float int_scale;
    float hbm_fraction;        /* share of free HBM to use when capacities are 0 */
} nh_server_config;

typedef struct {
    uint64_t submit_ns, admit_ns, done_ns;  /* steady-clock stamps */
    uint64_t t0;                            /* arena step of admission */
    uint64_t spikes;
    uint32_t n;                             /* neurons of the request */
} nh_request_stats;

typedef struct {
    uint64_t steps, events, spikes;       /* events/spikes of completed requests */
    uint64_t admitted, completed, peak_active;
    uint64_t cap_neurons, cap_synapses;
    double gpu_ms;
    uint64_t paired_rounds;               /* rounds run with two steps per launch */
} nh_server_stats;

typedef struct nh_server nh_server;
void nh_server_config_default(nh_server_config *c);
int nh_server_create(const nh_server_config *cfg, nh_server **out);
void nh_server_destroy(nh_server *s);
/* Copies everything it needs; stim is [stim_steps][n_inputs] (0 past it). */
int nh_server_submit(nh_server *s, const nh_net_desc *net, const nh_config *cfg, const float *stim,
                     uint64_t stim_steps, uint64_t total_steps, uint32_t win_steps, uint64_t *id);
/* Blocks up to timeout_ms; counts is [n_windows][n]. NH_ERR_NONE on timeout. */
int nh_server_wait(nh_server *s, uint64_t id, uint32_t timeout_ms, uint32_t *counts, uint64_t cap,
                   nh_request_stats *st);
int nh_server_get_stats(nh_server *s, nh_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif
