/*
 * oracle.h -- CPU restatement of the reference simulator, used ONLY as a
 * test checker and as the timed CPU baseline.  It is never linked into the
 * product library and nothing on the product path calls it.
 *
 * What it restates (reference files, relative to the reference root):
 *   sim/src/cpu.rs:192-326       CpuSim::step / run_steps (update order)
 *   sim/src/cpu.rs:162-184       reset_state
 *   sim/src/cpu.rs:409-486       run_ms, spike_row, spike_counts
 *   sim/src/state.rs:7-49        SimConfig defaults
 *   sim/src/state.rs:98-153      NeuronParams::from_network (model table, padding)
 *   sim/src/state.rs:221-224     ring_depth
 *   sim/src/state.rs:245-274     XSyn::build (custom-tau synapses)
 *   sim/src/state.rs:283-301     noise_hash / noise_fires
 *   sim/src/model/izh.rs:300-317 Izhikevich init + step
 *   sim/src/model/adex.rs:149-191 AdEx init + step + noise gain
 *   sim/tests/lif.rs:14-104      the LIF test model (init, step; gain 1)
 *   sim/src/kernels.rs:57-66     int ring put/get (fixed point, lrintf)
 *   sim/src/stim.rs:34-127       StimMap / Stimulus::value
 *   net-format/src/rng.rs:17-81  xoshiro256** Rng
 *   net-format/src/network.rs:451-692  add_neuron, set_kind, connect,
 *                                conduction_delay_ms, synapse_csr
 *   net-format/src/builders.rs:11-296  grids, create_connections, exc/inh,
 *                                set_synapse_tau_dist
 *   sim/src/bench.rs:10-53       bench_net
 *
 * Additions that the reference does not have (defined identically in the
 * product so both sides can be compared bit for bit):
 *   - int ring on the CPU (the reference has it on CUDA only) with the scale
 *     as a parameter (4096 in the reference, 65536 = Q16.16 also used);
 *   - Gaussian background current "noise_sd" (Irwin-Hall over 12 x 8-bit
 *     fields: the low half of the kick hash and one more splitmix64 hash);
 *   - exp_mode 1: a portable exp (Cody-Waite reduction + degree-7 Taylor
 *     polynomial, plain f32 operations) for AdEx, so CPU and GPU agree
 *     exactly; exp_mode 0 calls libm expf as the reference CPU backend does;
 *   - an FP64 instance of the same step, for stating float-mode error.
 */
#ifndef NEURO_ORACLE_H
#define NEURO_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- network descriptor (flat arrays, reference CSR order) ---- */
typedef struct {
    uint32_t n;
    uint64_t n_syn;
    const uint32_t *src_ptr;   /* n+1 */
    const uint32_t *tgt;       /* n_syn */
    const float *w;            /* n_syn, signed (Dale applied at CSR build) */
    const uint16_t *delay;     /* n_syn, steps >= 1 */
    const float *tau;          /* n_syn or NULL (all 0 = merged path) */
    uint32_t n_models;         /* entries in model_ids (<= 8) */
    const uint8_t *model_ids;  /* registry id per table entry: 0 izh, 1 adex, 2 lif */
    const uint8_t *model;      /* n: index into the table */
    uint32_t n_params, n_state;
    const float *p;            /* n * n_params */
    const float *x0;           /* n * n_state */
    uint32_t n_inputs;
    const uint32_t *stim_ptr;  /* n+1 or NULL */
    const uint32_t *stim_input;
    const float *stim_w;
} or_net_t;

typedef struct {
    float dt_ms;
    float synapse_tau_ms;
    float readout_tau_ms;
    float noise_rate_hz;
    float noise_kick_mv;
    float noise_sd;            /* Gaussian current sd per step (0 = off) */
    uint64_t seed;
    int32_t record_spikes;
    uint32_t spike_record_steps;
    int32_t int_weights;
    float int_scale;           /* 4096 = reference; 65536 = Q16.16 */
    int32_t exp_mode;          /* 0 libm expf, 1 portable exp */
    int32_t precision;         /* 32 or 64 */
} or_cfg_t;

void or_cfg_default(or_cfg_t *c);

/* ---- simulator ---- */
typedef struct or_sim or_sim;
or_sim *or_sim_new(const or_net_t *net, const or_cfg_t *cfg, char *err, int errlen);
void or_sim_free(or_sim *s);
void or_sim_reset(or_sim *s);
/* Run `steps` steps.  If cnt_out != NULL, cnt_out[k] = spikes at step k.  If
 * ids_out != NULL, the spiking ids of every step are appended in ascending
 * id order (at most ids_cap; *n_ids gets the number that would have been
 * written).  Returns 0, or -1 when ids_cap was too small. */
int or_sim_run(or_sim *s, uint64_t steps, uint32_t *cnt_out, uint32_t *ids_out,
               uint64_t ids_cap, uint64_t *n_ids);
uint64_t or_sim_t(const or_sim *s);
uint64_t or_sim_spikes_total(const or_sim *s);
uint32_t or_sim_ring_depth(const or_sim *s);
/* measurement counter: spike x out-synapse deliveries so far */
uint64_t or_sim_events_total(const or_sim *s);
void or_sim_state(const or_sim *s, uint32_t slot, double *out);  /* n */
void or_sim_syn_current(const or_sim *s, double *out);           /* s[i] */
void or_sim_trace(const or_sim *s, double *out);
void or_sim_pn_total(const or_sim *s, uint64_t *out);
void or_sim_xsyn(const or_sim *s, double *out, uint64_t cap);
uint64_t or_sim_n_xsyn(const or_sim *s);
/* spike_row: returns 1 and fills words (ceil(n/64)) or 0 (None). */
int or_sim_spike_row(const or_sim *s, uint64_t back, uint64_t *out);
/* spike_counts: 0 ok, -1 "window exceeds the recorded steps" */
int or_sim_spike_counts(const or_sim *s, uint64_t rows, uint32_t *out);
void or_sim_set_external(or_sim *s, uint32_t i, float amps);
void or_sim_set_input_constant(or_sim *s, uint32_t k, float v);
void or_sim_play_input(or_sim *s, uint32_t k, const float *series, uint64_t len, int looped);
void or_sim_set_noise(or_sim *s, float rate_hz, float kick_mv);
void or_sim_set_grid_watches(or_sim *s, uint32_t n_groups, const uint32_t *sizes,
                             const uint32_t *members);
/* out: concatenated texel traces for every group */
void or_sim_grid_traces(const or_sim *s, float tau_ms, float *out);

/* ---- shared arithmetic pieces (exported for known-answer tests) ---- */
uint32_t or_noise_hash(uint32_t i, uint64_t t, uint64_t seed);
int or_noise_fires(uint32_t i, uint64_t t, uint64_t seed, float p);
float or_gauss_draw(uint32_t i, uint64_t t, uint64_t seed);
float or_portable_expf(float x);
int32_t or_ring_put(float w, float scale);

/* ---- net-format restatement ---- */
typedef struct or_rng { uint64_t s[4]; } or_rng;
void or_rng_init(or_rng *r, uint64_t seed);
uint64_t or_rng_next_u64(or_rng *r);
double or_rng_f64(or_rng *r);
float or_rng_f32(or_rng *r);
double or_rng_uniform(or_rng *r, double lo, double hi);
uint64_t or_rng_below(or_rng *r, uint64_t n);
double or_rng_gauss(or_rng *r, double mean, double sd);

typedef struct or_netb or_netb;
or_netb *or_netb_new(void);
void or_netb_free(or_netb *b);
void or_netb_set_speed(or_netb *b, float speed);
/* kind: 0 excitatory, 1 inhibitory, 2 regular */
uint32_t or_netb_add_neuron(or_netb *b, float x, float y, float z, int kind);
void or_netb_set_kind(or_netb *b, uint32_t id, int kind);
/* model: 0 izh (4 params), 1 adex (10 params), 2 lif (5 params) */
void or_netb_set_model(or_netb *b, uint32_t id, int model, const float *params);
uint64_t or_netb_connect(or_netb *b, uint32_t src, uint32_t tgt, float w);
void or_netb_push_synapse(or_netb *b, uint32_t src, uint32_t tgt, float w, float delay_ms,
                          float tau_ms);
/* builders return the index of the group they created */
uint32_t or_netb_grid2d(or_netb *b, uint32_t width, uint32_t height, float dist,
                        float cx, float cy, float cz);
uint32_t or_netb_grid3d(or_netb *b, uint32_t width, uint32_t height, uint32_t depth,
                        float dist, float cx, float cy, float cz);
void or_netb_create_connections(or_netb *b, uint32_t group, double scaling, double lambda,
                                float weight, or_rng *r);
void or_netb_set_exc_inh(or_netb *b, uint32_t group, double inh_fraction, or_rng *r);
void or_netb_set_synapse_tau_dist(or_netb *b, uint32_t src_group, uint32_t tgt_group, double mean,
                                  double sd, or_rng *r);
void or_netb_set_exc_inh_exact(or_netb *b, uint32_t group, uint64_t n_inh, or_rng *r);
void or_netb_bench_net(or_netb *b, uint32_t n, uint32_t k, uint64_t seed, float w_scale,
                       float w_inh);
uint32_t or_netb_add_input(or_netb *b, uint32_t n_targets, const uint32_t *tgt,
                           const float *w);
uint32_t or_netb_n(const or_netb *b);
uint64_t or_netb_n_syn(const or_netb *b);
/* Build CSR + params into caller buffers sized by or_netb_n / n_syn.
 * Returns 0 or -1 with err set (delay over the u16 ring). */
int or_netb_csr(const or_netb *b, float dt_ms, uint32_t *src_ptr, uint32_t *tgt, float *w,
                uint16_t *delay, float *tau, char *err, int errlen);
/* model table: n_models/model_ids (<= 8), per-neuron model idx, padded
 * params and x0 (widths returned in *n_params / *n_state). p and x0 must
 * hold n*10 and n*2 floats. */
void or_netb_params(const or_netb *b, uint32_t *n_models, uint8_t *model_ids, uint8_t *model,
                    uint32_t *n_params, uint32_t *n_state, float *p, float *x0);
uint32_t or_netb_n_inputs(const or_netb *b);
uint64_t or_netb_n_stim(const or_netb *b);
void or_netb_stim_map(const or_netb *b, uint32_t *stim_ptr, uint32_t *stim_input,
                      float *stim_w);
void or_netb_positions(const or_netb *b, float *xyz);
void or_netb_kinds(const or_netb *b, uint8_t *kind);

#ifdef __cplusplus
}
#endif
#endif
