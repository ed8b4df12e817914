/*
 * sim_impl.h -- one precision instance of the oracle simulator.  Included by
 * sim.c twice: REAL=float (SFX f32) and REAL=double (SFX f64).
 * Test infrastructure only (see oracle.h).
 *
 * Mirrors CpuSim (sim/src/cpu.rs:20-306) field for field where the
 * reference has a field, in the same update order.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define SIMT CAT(or_simimpl_, SFX)
#define FN(name) CAT(name##_, SFX)

typedef struct {
    uint32_t n, d, words, np, nx, n_inputs, n_models;
    uint8_t model_ids[8];
    or_cfg_t cfg;
    /* network */
    uint32_t *src_ptr, *tgt;
    REAL *w;       /* float ring: signed weight */
    int32_t *wq;   /* int ring: quantised weight */
    uint16_t *delay;
    uint8_t *model;
    REAL *p;
    REAL *x0;
    uint32_t *stim_ptr, *stim_input;
    REAL *stim_w;
    /* state */
    REAL *x, *s, *trace, *ring;
    int32_t *iring;
    REAL *ext;
    float *ext_f;
    /* custom-tau synapses (sim/src/state.rs:245-274) */
    uint64_t n_x;
    uint32_t *xmap;
    REAL *xdecay, *xw, *xsyn;
    uint32_t *xt_ptr, *xt_idx;
    uint32_t **xring;
    uint64_t *xring_len, *xring_cap;
    /* stimulus program (sim/src/stim.rs:9-21) */
    float *constants;
    float **series;
    uint64_t *series_len, *play_start;
    int *looped;
    /* recording */
    uint32_t R;
    uint64_t *spike_bits;
    uint32_t *spike_counts;
    uint64_t *pn_total;
    /* grid watches (sim/src/cpu.rs:85-92) */
    uint32_t n_groups;
    uint32_t *g_size;
    int32_t **g_map;
    uint32_t **g_ls;
    uint64_t t, spikes_total;
    uint64_t events_total;  /* measurement only: spike x out-synapse deliveries */
    uint32_t *spiked;
} SIMT;

static void FN(sim_free)(SIMT *S)
{
    if (!S)
        return;
    free(S->src_ptr); free(S->tgt); free(S->w); free(S->wq); free(S->delay);
    free(S->model); free(S->p); free(S->x0); free(S->stim_ptr); free(S->stim_input);
    free(S->stim_w); free(S->x); free(S->s); free(S->trace); free(S->ring); free(S->iring);
    free(S->ext); free(S->ext_f); free(S->xmap); free(S->xdecay); free(S->xw); free(S->xsyn);
    free(S->xt_ptr); free(S->xt_idx);
    if (S->xring)
        for (uint32_t k = 0; k < S->d; ++k)
            free(S->xring[k]);
    free(S->xring); free(S->xring_len); free(S->xring_cap);
    free(S->constants);
    if (S->series)
        for (uint32_t k = 0; k < S->n_inputs; ++k)
            free(S->series[k]);
    free(S->series); free(S->series_len); free(S->play_start); free(S->looped);
    free(S->spike_bits); free(S->spike_counts); free(S->pn_total);
    for (uint32_t g = 0; g < S->n_groups; ++g) {
        free(S->g_map[g]);
        free(S->g_ls[g]);
    }
    free(S->g_size); free(S->g_map); free(S->g_ls);
    free(S->spiked);
    free(S);
}

/* sim/src/cpu.rs:162-184 */
static void FN(sim_reset)(SIMT *S)
{
    uint64_t n = S->n;
    memcpy(S->x, S->x0, sizeof(REAL) * n * S->nx);
    memset(S->s, 0, sizeof(REAL) * n);
    memset(S->trace, 0, sizeof(REAL) * n);
    if (S->ring)
        memset(S->ring, 0, sizeof(REAL) * n * S->d);
    if (S->iring)
        memset(S->iring, 0, sizeof(int32_t) * n * S->d);
    if (S->n_x)
        memset(S->xsyn, 0, sizeof(REAL) * S->n_x);
    for (uint32_t k = 0; k < S->d; ++k)
        S->xring_len[k] = 0;
    if (S->spike_bits)
        memset(S->spike_bits, 0, sizeof(uint64_t) * (uint64_t)S->R * S->words);
    if (S->spike_counts)
        memset(S->spike_counts, 0, sizeof(uint32_t) * S->R);
    memset(S->pn_total, 0, sizeof(uint64_t) * n);
    S->t = 0;
    S->spikes_total = 0;
    S->events_total = 0;
    for (uint32_t g = 0; g < S->n_groups; ++g)
        for (uint32_t k = 0; k < S->g_size[g]; ++k)
            S->g_ls[g][k] = UINT32_MAX;
}

static SIMT *FN(sim_new)(const or_net_t *net, const or_cfg_t *cfg, char *err, int errlen)
{
    SIMT *S = (SIMT *)calloc(1, sizeof(SIMT));
    uint64_t n = net->n, E = net->n_syn;
    S->n = net->n;
    S->cfg = *cfg;
    S->np = net->n_params;
    S->nx = net->n_state;
    S->n_models = net->n_models;
    if (net->n_models > 8) {
        snprintf(err, errlen, "model table too large");
        free(S);
        return NULL;
    }
    for (uint32_t k = 0; k < net->n_models; ++k)
        S->model_ids[k] = net->model_ids[k];
    for (uint64_t i = 0; i < n; ++i)
        if (net->model[i] >= net->n_models) {
            snprintf(err, errlen, "neuron %llu names model %u outside the model table",
                     (unsigned long long)i, net->model[i]);
            free(S);
            return NULL;
        }
    S->words = (uint32_t)((n + 63) / 64);
    /* ring_depth: sim/src/state.rs:221-224 */
    uint32_t maxd = 1;
    for (uint64_t j = 0; j < E; ++j)
        if (net->delay[j] > maxd)
            maxd = net->delay[j];
    if (E == 0)
        maxd = 1;
    S->d = maxd + 1 < 2 ? 2 : maxd + 1;
    S->src_ptr = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    memcpy(S->src_ptr, net->src_ptr, sizeof(uint32_t) * (n + 1));
    S->tgt = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    memcpy(S->tgt, net->tgt, sizeof(uint32_t) * E);
    S->delay = (uint16_t *)malloc(sizeof(uint16_t) * (E ? E : 1));
    memcpy(S->delay, net->delay, sizeof(uint16_t) * E);
    S->w = (REAL *)malloc(sizeof(REAL) * (E ? E : 1));
    for (uint64_t j = 0; j < E; ++j)
        S->w[j] = (REAL)net->w[j];
    if (cfg->int_weights) {
        S->wq = (int32_t *)malloc(sizeof(int32_t) * (E ? E : 1));
        for (uint64_t j = 0; j < E; ++j)
            S->wq[j] = or_ring_put(net->w[j], cfg->int_scale);
        S->iring = (int32_t *)calloc(n * S->d + 1, sizeof(int32_t));
    } else {
        S->ring = (REAL *)calloc(n * S->d + 1, sizeof(REAL));
    }
    S->model = (uint8_t *)malloc(n + 1);
    memcpy(S->model, net->model, n);
    S->p = (REAL *)malloc(sizeof(REAL) * (n * S->np + 1));
    for (uint64_t k = 0; k < n * S->np; ++k)
        S->p[k] = (REAL)net->p[k];
    S->x0 = (REAL *)malloc(sizeof(REAL) * (n * S->nx + 1));
    for (uint64_t k = 0; k < n * S->nx; ++k)
        S->x0[k] = (REAL)net->x0[k];
    S->n_inputs = net->n_inputs;
    S->stim_ptr = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    uint64_t ns = 0;
    if (net->stim_ptr) {
        memcpy(S->stim_ptr, net->stim_ptr, sizeof(uint32_t) * (n + 1));
        ns = net->stim_ptr[n];
    }
    S->stim_input = (uint32_t *)malloc(sizeof(uint32_t) * (ns + 1));
    S->stim_w = (REAL *)malloc(sizeof(REAL) * (ns + 1));
    for (uint64_t j = 0; j < ns; ++j) {
        S->stim_input[j] = net->stim_input[j];
        S->stim_w[j] = (REAL)net->stim_w[j];
    }
    S->x = (REAL *)malloc(sizeof(REAL) * (n * S->nx + 1));
    S->s = (REAL *)calloc(n + 1, sizeof(REAL));
    S->trace = (REAL *)calloc(n + 1, sizeof(REAL));
    S->ext = (REAL *)calloc(n + 1, sizeof(REAL));
    S->ext_f = (float *)calloc(n + 1, sizeof(float));
    /* XSyn::build: sim/src/state.rs:245-274 */
    S->xmap = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    uint64_t nxs = 0;
    for (uint64_t j = 0; j < E; ++j)
        S->xmap[j] = (net->tau && net->tau[j] > 0.0f) ? (uint32_t)nxs++ : UINT32_MAX;
    S->n_x = nxs;
    S->xdecay = (REAL *)malloc(sizeof(REAL) * (nxs + 1));
    S->xw = (REAL *)malloc(sizeof(REAL) * (nxs + 1));
    S->xsyn = (REAL *)calloc(nxs + 1, sizeof(REAL));
    S->xt_ptr = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    S->xt_idx = (uint32_t *)malloc(sizeof(uint32_t) * (nxs + 1));
    if (nxs) {
        uint32_t *xtgt = (uint32_t *)malloc(sizeof(uint32_t) * nxs);
        for (uint64_t j = 0; j < E; ++j) {
            uint32_t k = S->xmap[j];
            if (k == UINT32_MAX)
                continue;
            float tau = net->tau[j] > 1e-6f ? net->tau[j] : 1e-6f;
            S->xdecay[k] = EXPR((REAL)(-cfg->dt_ms) / (REAL)tau);
            if (cfg->int_weights)
                /* CUDA int ring: the arrival adds rt_get(rt_put(w)) */
                S->xw[k] = (REAL)S->wq[j] / (REAL)cfg->int_scale;
            else
                S->xw[k] = (REAL)net->w[j];
            xtgt[k] = net->tgt[j];
            S->xt_ptr[net->tgt[j] + 1]++;
        }
        for (uint64_t i = 0; i < n; ++i)
            S->xt_ptr[i + 1] += S->xt_ptr[i];
        uint32_t *fill = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
        for (uint64_t k = 0; k < nxs; ++k) {
            uint32_t tg = xtgt[k];
            S->xt_idx[S->xt_ptr[tg] + fill[tg]++] = (uint32_t)k;
        }
        free(fill);
        free(xtgt);
    }
    S->xring = (uint32_t **)calloc(S->d, sizeof(uint32_t *));
    S->xring_len = (uint64_t *)calloc(S->d, sizeof(uint64_t));
    S->xring_cap = (uint64_t *)calloc(S->d, sizeof(uint64_t));
    /* stimulus */
    uint32_t ni = S->n_inputs ? S->n_inputs : 1;
    S->constants = (float *)calloc(ni, sizeof(float));
    S->series = (float **)calloc(ni, sizeof(float *));
    S->series_len = (uint64_t *)calloc(ni, sizeof(uint64_t));
    S->play_start = (uint64_t *)calloc(ni, sizeof(uint64_t));
    S->looped = (int *)calloc(ni, sizeof(int));
    /* recording: spike_record_steps.max(2) when on (cpu.rs:112-116) */
    S->R = cfg->record_spikes ? (cfg->spike_record_steps < 2 ? 2 : cfg->spike_record_steps) : 0;
    if (S->R) {
        S->spike_bits = (uint64_t *)calloc((uint64_t)S->R * S->words + 1, sizeof(uint64_t));
        S->spike_counts = (uint32_t *)calloc(S->R, sizeof(uint32_t));
    }
    S->pn_total = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    S->spiked = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    FN(sim_reset)(S);
    (void)err;
    (void)errlen;
    return S;
}

/* Stimulus::value -- sim/src/stim.rs:102-115 */
static float FN(stim_value)(const SIMT *S, uint32_t k, uint64_t t)
{
    if (k < S->n_inputs && S->series_len[k] > 0) {
        uint64_t len = S->series_len[k];
        uint64_t start = S->play_start[k];
        if (S->looped[k]) {
            int64_t d = (int64_t)t - (int64_t)start;
            int64_t m = d % (int64_t)len;
            if (m < 0)
                m += (int64_t)len;
            return S->series[k][m];
        }
        uint64_t idx = t - start;
        return idx < len ? S->series[k][idx] : 0.0f;
    }
    return k < S->n_inputs ? S->constants[k] : 0.0f;
}

static void FN(xring_push)(SIMT *S, uint32_t slot, uint32_t id)
{
    if (S->xring_len[slot] == S->xring_cap[slot]) {
        uint64_t c = S->xring_cap[slot] ? 2 * S->xring_cap[slot] : 16;
        S->xring[slot] = (uint32_t *)realloc(S->xring[slot], sizeof(uint32_t) * c);
        S->xring_cap[slot] = c;
    }
    S->xring[slot][S->xring_len[slot]++] = id;
}

/* CpuSim::step -- sim/src/cpu.rs:192-306, same order. */
static uint32_t FN(sim_step)(SIMT *S)
{
    const uint64_t n = S->n;
    const float dtf = S->cfg.dt_ms;
    const REAL dt = (REAL)dtf;
    const uint64_t t = S->t;
    const uint32_t slot = (uint32_t)(t % S->d);
    /* cpu.rs:197-204: kick probability and size in f32; decays via exp */
    const float p_kick = S->cfg.noise_rate_hz * dtf / 1000.0f;
    const REAL kick = S->cfg.noise_kick_mv != 0.0f ? (REAL)S->cfg.noise_kick_mv / dt : LIT(0.0);
    const float st = S->cfg.synapse_tau_ms > 1e-6f ? S->cfg.synapse_tau_ms : 1e-6f;
    const float rt = S->cfg.readout_tau_ms > 1e-6f ? S->cfg.readout_tau_ms : 1e-6f;
    const REAL decay_s = EXPR(-dt / (REAL)st);
    const REAL decay_t = EXPR(-dt / (REAL)rt);
    const REAL sd = (REAL)S->cfg.noise_sd;

    /* custom-tau arrivals (cpu.rs:208-211) */
    for (uint64_t a = 0; a < S->xring_len[slot]; ++a) {
        uint32_t k = S->xring[slot][a];
        S->xsyn[k] += S->xw[k];
    }
    S->xring_len[slot] = 0;

    const uint64_t base = (uint64_t)slot * n;
    uint32_t ns = 0;
    for (uint64_t i = 0; i < n; ++i) {
        REAL arr;
        if (S->iring) {
            arr = (REAL)S->iring[base + i] / (REAL)S->cfg.int_scale;
            S->iring[base + i] = 0;
        } else {
            arr = S->ring[base + i];
            S->ring[base + i] = LIT(0.0);
        }
        S->s[i] += arr;
        REAL input = S->s[i] + S->ext[i];
        for (uint32_t j = S->xt_ptr[i]; j < S->xt_ptr[i + 1]; ++j) {
            uint32_t k = S->xt_idx[j];
            input += S->xsyn[k];
            S->xsyn[k] *= S->xdecay[k];
        }
        for (uint32_t j = S->stim_ptr[i]; j < S->stim_ptr[i + 1]; ++j)
            input += S->stim_w[j] * (REAL)FN(stim_value)(S, S->stim_input[j], t);
        REAL *p = S->p + i * S->np;
        REAL *x = S->x + i * S->nx;
        const uint8_t mid = S->model_ids[S->model[i]];
        const REAL gain = mid == 1 ? p[0] : LIT(1.0);
        if (p_kick > 0.0f && or_noise_fires((uint32_t)i, t, S->cfg.seed, p_kick))
            input += kick * gain;
        if (S->cfg.noise_sd != 0.0f)
            input += (sd * (REAL)or_gauss_draw((uint32_t)i, t, S->cfg.seed)) * gain;
        int sp;
        if (mid == 0) {
            /* model/izh.rs:309-317 */
            if (x[0] >= LIT(30.0)) {
                x[0] = p[2];
                x[1] += p[3];
            }
            x[0] += dt * (LIT(0.04) * x[0] * x[0] + LIT(5.0) * x[0] + LIT(140.0) - x[1] + input);
            x[1] += dt * p[0] * (p[1] * x[0] - x[1]);
            sp = x[0] >= LIT(30.0);
        } else if (mid == 2) {
            /* the LIF test model (sim/tests/lif.rs:98-104): reset for last
             * step's crossing, v += dt*(e - v + r*I)/tau, spike test */
            if (x[0] >= p[3])
                x[0] = p[4];
            x[0] += dt * (p[1] - x[0] + p[2] * input) / p[0];
            sp = x[0] >= p[3];
        } else {
            /* model/adex.rs:163-176 */
            if (x[0] >= p[9]) {
                x[0] = p[8];
                x[1] += p[7];
            }
            REAL q = (x[0] - p[3]) / p[4];
            if (q > LIT(10.0))
                q = LIT(10.0);
            REAL ex = (S->cfg.exp_mode == 1 && !IS_F64) ? (REAL)or_portable_expf((float)q) : EXPR(q);
            x[0] += dt * (-p[1] * (x[0] - p[2]) + p[1] * p[4] * ex - x[1] + input) / p[0];
            x[1] += dt * (p[5] * (x[0] - p[2]) - x[1]) / p[6];
            sp = x[0] >= p[9];
            if (x[0] > p[9])
                x[0] = p[9];
        }
        if (sp) {
            S->spiked[ns++] = (uint32_t)i;
            S->trace[i] = LIT(1.0);
        } else {
            S->trace[i] *= decay_t;
        }
        S->s[i] *= decay_s;
    }

    /* scatter (cpu.rs:264-278) */
    for (uint32_t a = 0; a < ns; ++a) {
        uint32_t i = S->spiked[a];
        S->events_total += S->src_ptr[i + 1] - S->src_ptr[i];
        for (uint32_t j = S->src_ptr[i]; j < S->src_ptr[i + 1]; ++j) {
            uint32_t dst = (uint32_t)((t + S->delay[j]) % S->d);
            uint32_t xid = S->xmap[j];
            if (xid != UINT32_MAX)
                FN(xring_push)(S, dst, xid);
            else if (S->iring)
                S->iring[(uint64_t)dst * n + S->tgt[j]] =
                    (int32_t)((uint32_t)S->iring[(uint64_t)dst * n + S->tgt[j]] + (uint32_t)S->wq[j]);
            else
                S->ring[(uint64_t)dst * n + S->tgt[j]] += S->w[j];
        }
    }

    /* recording (cpu.rs:280-289) */
    if (S->R) {
        uint64_t row = t % S->R;
        S->spike_counts[row] = ns;
        uint64_t *bits = S->spike_bits + row * S->words;
        memset(bits, 0, sizeof(uint64_t) * S->words);
        for (uint32_t a = 0; a < ns; ++a)
            bits[S->spiked[a] / 64] |= 1ull << (S->spiked[a] % 64);
    }
    for (uint32_t a = 0; a < ns; ++a)
        S->pn_total[S->spiked[a]] += 1;
    if (S->n_groups) {
        uint32_t t32 = (uint32_t)(t < (uint64_t)UINT32_MAX - 1 ? t : (uint64_t)UINT32_MAX - 1);
        for (uint32_t a = 0; a < ns; ++a)
            for (uint32_t g = 0; g < S->n_groups; ++g) {
                int32_t k = S->g_map[g][S->spiked[a]];
                if (k >= 0)
                    S->g_ls[g][k] = t32;
            }
    }
    S->spikes_total += ns;
    S->t += 1;
    return ns;
}

static int FN(sim_run)(SIMT *S, uint64_t steps, uint32_t *cnt_out, uint32_t *ids_out,
                       uint64_t ids_cap, uint64_t *n_ids)
{
    uint64_t w = 0;
    for (uint64_t k = 0; k < steps; ++k) {
        uint32_t ns = FN(sim_step)(S);
        if (cnt_out)
            cnt_out[k] = ns;
        if (ids_out) {
            for (uint32_t a = 0; a < ns; ++a) {
                if (w < ids_cap)
                    ids_out[w] = S->spiked[a];
                ++w;
            }
        }
    }
    if (n_ids)
        *n_ids = w;
    return (ids_out && w > ids_cap) ? -1 : 0;
}

/* CpuSim::spike_row -- cpu.rs:429-440 */
static int FN(sim_spike_row)(const SIMT *S, uint64_t back, uint64_t *out)
{
    if (!S->cfg.record_spikes || S->t == 0 || back >= S->t || back >= S->R)
        return 0;
    uint64_t tt = S->t - 1 - back;
    uint64_t row = tt % S->R;
    memcpy(out, S->spike_bits + row * S->words, sizeof(uint64_t) * S->words);
    return 1;
}

/* CpuSim::spike_counts -- cpu.rs:452-486 */
static int FN(sim_spike_counts)(const SIMT *S, uint64_t rows, uint32_t *out)
{
    memset(out, 0, sizeof(uint32_t) * S->n);
    if (!S->cfg.record_spikes || S->t == 0)
        return 0;
    if (rows >= S->t) {
        for (uint32_t i = 0; i < S->n; ++i)
            out[i] = S->pn_total[i] > UINT32_MAX ? UINT32_MAX : (uint32_t)S->pn_total[i];
        return 0;
    }
    if (rows > S->R)
        return -1;
    for (uint64_t back = 0; back < rows; ++back) {
        if (back >= S->t)
            break;
        uint64_t row = (S->t - 1 - back) % S->R;
        const uint64_t *bits = S->spike_bits + row * S->words;
        for (uint32_t wi = 0; wi < S->words; ++wi) {
            uint64_t v = bits[wi];
            while (v) {
                uint32_t b = (uint32_t)__builtin_ctzll(v);
                v &= v - 1;
                uint64_t i = (uint64_t)wi * 64 + b;
                if (i < S->n)
                    out[i] += 1;
            }
        }
    }
    return 0;
}

/* CpuSim::set_grid_watches -- cpu.rs:361-387 (the oracle rebuilds always) */
static void FN(sim_set_grid_watches)(SIMT *S, uint32_t n_groups, const uint32_t *sizes,
                                     const uint32_t *members)
{
    for (uint32_t g = 0; g < S->n_groups; ++g) {
        free(S->g_map[g]);
        free(S->g_ls[g]);
    }
    free(S->g_size); free(S->g_map); free(S->g_ls);
    S->n_groups = n_groups;
    S->g_size = (uint32_t *)calloc(n_groups + 1, sizeof(uint32_t));
    S->g_map = (int32_t **)calloc(n_groups + 1, sizeof(int32_t *));
    S->g_ls = (uint32_t **)calloc(n_groups + 1, sizeof(uint32_t *));
    uint64_t off = 0;
    for (uint32_t g = 0; g < n_groups; ++g) {
        S->g_size[g] = sizes[g];
        S->g_map[g] = (int32_t *)malloc(sizeof(int32_t) * (S->n + 1));
        for (uint32_t i = 0; i < S->n; ++i)
            S->g_map[g][i] = -1;
        for (uint32_t k = 0; k < sizes[g]; ++k) {
            uint32_t m = members[off + k];
            if (m < S->n)
                S->g_map[g][m] = (int32_t)k;
        }
        S->g_ls[g] = (uint32_t *)malloc(sizeof(uint32_t) * (sizes[g] + 1));
        for (uint32_t k = 0; k < sizes[g]; ++k)
            S->g_ls[g][k] = UINT32_MAX;
        off += sizes[g];
    }
}

/* CpuSim::grid_traces -- cpu.rs:391-407 (f32 expf, as the reference) */
static void FN(sim_grid_traces)(const SIMT *S, float tau_ms, float *out)
{
    float k = S->cfg.dt_ms / (tau_ms > 1e-6f ? tau_ms : 1e-6f);
    uint64_t o = 0;
    for (uint32_t g = 0; g < S->n_groups; ++g)
        for (uint32_t q = 0; q < S->g_size[g]; ++q) {
            uint32_t l = S->g_ls[g][q];
            if (l == UINT32_MAX) {
                out[o++] = 0.0f;
            } else {
                uint64_t gap = S->t > l ? S->t - l : 0;
                out[o++] = expf(-((float)gap) * k);
            }
        }
}

#undef SIMT
#undef FN
#undef CAT
#undef CAT2
