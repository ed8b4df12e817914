/*
 * sim.c -- oracle simulator front end: f32 and f64 instances of
 * sim_impl.h behind one handle.  Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define REAL float
#define SFX f32
#define LIT(x) x##f
#define EXPR(x) expf(x)
#define IS_F64 0
#include "sim_impl.h"
#undef REAL
#undef SFX
#undef LIT
#undef EXPR
#undef IS_F64

#define REAL double
#define SFX f64
#define LIT(x) x
#define EXPR(x) exp(x)
#define IS_F64 1
#include "sim_impl.h"
#undef REAL
#undef SFX
#undef LIT
#undef EXPR
#undef IS_F64

struct or_sim {
    int prec;
    or_simimpl_f32 *a;
    or_simimpl_f64 *b;
};

/* SimConfig::default -- sim/src/state.rs:35-49 (dt 0.1, tau 2/2, kick 20,
 * seed 1, 512 rows), plus the additions' neutral values. */
void or_cfg_default(or_cfg_t *c)
{
    memset(c, 0, sizeof(*c));
    c->dt_ms = 0.1f;
    c->synapse_tau_ms = 2.0f;
    c->readout_tau_ms = 2.0f;
    c->noise_rate_hz = 0.0f;
    c->noise_kick_mv = 20.0f;
    c->noise_sd = 0.0f;
    c->seed = 1;
    c->record_spikes = 0;
    c->spike_record_steps = 512;
    c->int_weights = 0;
    c->int_scale = 4096.0f;
    c->exp_mode = 0;
    c->precision = 32;
}

or_sim *or_sim_new(const or_net_t *net, const or_cfg_t *cfg, char *err, int errlen)
{
    or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
    s->prec = cfg->precision == 64 ? 64 : 32;
    if (s->prec == 32)
        s->a = sim_new_f32(net, cfg, err, errlen);
    else
        s->b = sim_new_f64(net, cfg, err, errlen);
    if (!s->a && !s->b) {
        free(s);
        return NULL;
    }
    return s;
}

void or_sim_free(or_sim *s)
{
    if (!s)
        return;
    sim_free_f32(s->a);
    sim_free_f64(s->b);
    free(s);
}

#define DISPATCH(call_a, call_b) \
    do {                         \
        if (s->a)                \
            call_a;              \
        else                     \
            call_b;              \
    } while (0)

void or_sim_reset(or_sim *s) { DISPATCH(sim_reset_f32(s->a), sim_reset_f64(s->b)); }

int or_sim_run(or_sim *s, uint64_t steps, uint32_t *cnt_out, uint32_t *ids_out,
               uint64_t ids_cap, uint64_t *n_ids)
{
    if (s->a)
        return sim_run_f32(s->a, steps, cnt_out, ids_out, ids_cap, n_ids);
    return sim_run_f64(s->b, steps, cnt_out, ids_out, ids_cap, n_ids);
}

uint64_t or_sim_t(const or_sim *s) { return s->a ? s->a->t : s->b->t; }
uint64_t or_sim_spikes_total(const or_sim *s)
{
    return s->a ? s->a->spikes_total : s->b->spikes_total;
}
uint32_t or_sim_ring_depth(const or_sim *s) { return s->a ? s->a->d : s->b->d; }
uint64_t or_sim_events_total(const or_sim *s)
{
    return s->a ? s->a->events_total : s->b->events_total;
}

void or_sim_state(const or_sim *s, uint32_t slot, double *out)
{
    uint32_t n = s->a ? s->a->n : s->b->n;
    uint32_t nx = s->a ? s->a->nx : s->b->nx;
    if (slot >= nx)
        return;
    for (uint32_t i = 0; i < n; ++i)
        out[i] = s->a ? (double)s->a->x[(uint64_t)i * nx + slot] : s->b->x[(uint64_t)i * nx + slot];
}

void or_sim_syn_current(const or_sim *s, double *out)
{
    uint32_t n = s->a ? s->a->n : s->b->n;
    for (uint32_t i = 0; i < n; ++i)
        out[i] = s->a ? (double)s->a->s[i] : s->b->s[i];
}

void or_sim_trace(const or_sim *s, double *out)
{
    uint32_t n = s->a ? s->a->n : s->b->n;
    for (uint32_t i = 0; i < n; ++i)
        out[i] = s->a ? (double)s->a->trace[i] : s->b->trace[i];
}

void or_sim_pn_total(const or_sim *s, uint64_t *out)
{
    uint32_t n = s->a ? s->a->n : s->b->n;
    memcpy(out, s->a ? s->a->pn_total : s->b->pn_total, sizeof(uint64_t) * n);
}

uint64_t or_sim_n_xsyn(const or_sim *s) { return s->a ? s->a->n_x : s->b->n_x; }

void or_sim_xsyn(const or_sim *s, double *out, uint64_t cap)
{
    uint64_t m = or_sim_n_xsyn(s);
    for (uint64_t k = 0; k < m && k < cap; ++k)
        out[k] = s->a ? (double)s->a->xsyn[k] : s->b->xsyn[k];
}

int or_sim_spike_row(const or_sim *s, uint64_t back, uint64_t *out)
{
    return s->a ? sim_spike_row_f32(s->a, back, out) : sim_spike_row_f64(s->b, back, out);
}

int or_sim_spike_counts(const or_sim *s, uint64_t rows, uint32_t *out)
{
    return s->a ? sim_spike_counts_f32(s->a, rows, out) : sim_spike_counts_f64(s->b, rows, out);
}

void or_sim_set_external(or_sim *s, uint32_t i, float amps)
{
    if (s->a) {
        if (i < s->a->n) {
            s->a->ext[i] = amps;
            s->a->ext_f[i] = amps;
        }
    } else if (i < s->b->n) {
        s->b->ext[i] = amps;
        s->b->ext_f[i] = amps;
    }
}

void or_sim_set_input_constant(or_sim *s, uint32_t k, float v)
{
    if (s->a) {
        if (k < s->a->n_inputs)
            s->a->constants[k] = v;
    } else if (k < s->b->n_inputs) {
        s->b->constants[k] = v;
    }
}

/* Stimulus::play / play_looped -- sim/src/stim.rs:72-95 (start = current t) */
void or_sim_play_input(or_sim *s, uint32_t k, const float *series, uint64_t len, int looped)
{
    uint32_t ni = s->a ? s->a->n_inputs : s->b->n_inputs;
    if (k >= ni)
        return;
    float **ser = s->a ? s->a->series : s->b->series;
    uint64_t *slen = s->a ? s->a->series_len : s->b->series_len;
    uint64_t *start = s->a ? s->a->play_start : s->b->play_start;
    int *lp = s->a ? s->a->looped : s->b->looped;
    free(ser[k]);
    ser[k] = (float *)malloc(sizeof(float) * (len ? len : 1));
    memcpy(ser[k], series, sizeof(float) * len);
    slen[k] = len;
    start[k] = or_sim_t(s);
    lp[k] = looped ? 1 : 0;
}

void or_sim_set_noise(or_sim *s, float rate_hz, float kick_mv)
{
    or_cfg_t *c = s->a ? &s->a->cfg : &s->b->cfg;
    c->noise_rate_hz = rate_hz;
    c->noise_kick_mv = kick_mv;
}

void or_sim_set_grid_watches(or_sim *s, uint32_t n_groups, const uint32_t *sizes,
                             const uint32_t *members)
{
    DISPATCH(sim_set_grid_watches_f32(s->a, n_groups, sizes, members),
             sim_set_grid_watches_f64(s->b, n_groups, sizes, members));
}

void or_sim_grid_traces(const or_sim *s, float tau_ms, float *out)
{
    if (s->a)
        sim_grid_traces_f32(s->a, tau_ms, out);
    else
        sim_grid_traces_f64(s->b, tau_ms, out);
}
