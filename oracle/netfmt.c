/*
 * netfmt.c -- restatement of the reference's net-format pieces that feed
 * the simulator: Rng, the editable network's neuron/type/synapse rules,
 * the builders, bench_net and synapse_csr.  Test infrastructure only (see
 * oracle.h).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ---------------- Rng: net-format/src/rng.rs:17-81 ---------------- */
static uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void or_rng_init(or_rng *r, uint64_t seed)
{
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t v = z;
        v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
        v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
        r->s[i] = v ^ (v >> 31);
    }
}

uint64_t or_rng_next_u64(or_rng *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

double or_rng_f64(or_rng *r) { return (double)(or_rng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
float or_rng_f32(or_rng *r) { return (float)(or_rng_next_u64(r) >> 40) * (1.0f / 16777216.0f); }
double or_rng_uniform(or_rng *r, double lo, double hi) { return lo + (hi - lo) * or_rng_f64(r); }
uint64_t or_rng_below(or_rng *r, uint64_t n) { return or_rng_next_u64(r) % (n ? n : 1); }

double or_rng_gauss(or_rng *r, double mean, double sd)
{
    double u1 = 1.0 - or_rng_f64(r);
    if (u1 < 1e-300)
        u1 = 1e-300;
    double u2 = or_rng_f64(r);
    return mean + sd * sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

/* ---------------- network (net-format/src/network.rs) ---------------- */
static const float IZH_RS[4] = {0.02f, 0.2f, -65.0f, 8.0f};
static const float IZH_FS[4] = {0.1f, 0.2f, -65.0f, 2.0f};

typedef struct {
    float pos[3];
    int kind;      /* 0 exc, 1 inh, 2 regular */
    int model;     /* registry id: 0 izh, 1 adex */
    int np;        /* effective param-row width */
    float p[10];
} nb_neuron;

typedef struct { uint32_t src, tgt; float w, delay_ms, tau_ms; } nb_syn;

typedef struct { uint32_t *m; uint32_t len; } nb_group;

typedef struct { uint32_t *tgt; float *w; uint32_t len; } nb_input;

struct or_netb {
    nb_neuron *nr; uint32_t n, ncap;
    nb_syn *sy; uint64_t ns, scap;
    nb_group *gr; uint32_t ng;
    nb_input *in; uint32_t ni;
    float speed;   /* Meta::conduction_speed, default 100 (lib.rs:36) */
};

or_netb *or_netb_new(void)
{
    or_netb *b = (or_netb *)calloc(1, sizeof(or_netb));
    b->speed = 100.0f;
    return b;
}

void or_netb_free(or_netb *b)
{
    if (!b)
        return;
    free(b->nr);
    free(b->sy);
    for (uint32_t g = 0; g < b->ng; ++g)
        free(b->gr[g].m);
    free(b->gr);
    for (uint32_t k = 0; k < b->ni; ++k) {
        free(b->in[k].tgt);
        free(b->in[k].w);
    }
    free(b->in);
    free(b);
}

void or_netb_set_speed(or_netb *b, float speed) { b->speed = speed; }
uint32_t or_netb_n(const or_netb *b) { return b->n; }
uint64_t or_netb_n_syn(const or_netb *b) { return b->ns; }

static void kind_default(int kind, nb_neuron *x)
{
    x->model = 0;
    x->np = 4;
    memcpy(x->p, kind == 1 ? IZH_FS : IZH_RS, sizeof(IZH_RS));
}

/* Network::add_neuron -- network.rs:451-461: the stock izh row of the kind */
uint32_t or_netb_add_neuron(or_netb *b, float x, float y, float z, int kind)
{
    if (b->n == b->ncap) {
        b->ncap = b->ncap ? 2 * b->ncap : 64;
        b->nr = (nb_neuron *)realloc(b->nr, sizeof(nb_neuron) * b->ncap);
    }
    nb_neuron *r = &b->nr[b->n];
    memset(r, 0, sizeof(*r));
    r->pos[0] = x;
    r->pos[1] = y;
    r->pos[2] = z;
    r->kind = kind;
    kind_default(kind, r);
    return b->n++;
}

/* Network::set_kind -- network.rs:530-541: a neuron still on a stock izh
 * row (RS or FS) takes the new kind's stock row; a custom row stays as an
 * override on the kind's izh type. */
void or_netb_set_kind(or_netb *b, uint32_t id, int kind)
{
    nb_neuron *r = &b->nr[id];
    int stock = r->np == 4 && (memcmp(r->p, IZH_RS, sizeof(IZH_RS)) == 0 ||
                               memcmp(r->p, IZH_FS, sizeof(IZH_FS)) == 0);
    r->kind = kind;
    if (stock) {
        kind_default(kind, r);
    } else {
        r->model = 0; /* the kind's type is an izh type; the override stays */
    }
}

void or_netb_set_model(or_netb *b, uint32_t id, int model, const float *params)
{
    nb_neuron *r = &b->nr[id];
    r->model = model;
    r->np = model == 1 ? 10 : (model == 2 ? 5 : 4);
    memset(r->p, 0, sizeof(r->p));
    memcpy(r->p, params, sizeof(float) * r->np);
}

static float pair_dist(const nb_neuron *a, const nb_neuron *c)
{
    float dx = a->pos[0] - c->pos[0], dy = a->pos[1] - c->pos[1], dz = a->pos[2] - c->pos[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

/* builders::set_synapse_tau_dist (net-format/src/builders.rs:239-258):
 * every synapse from a member of group sg to a member of group tg, in
 * insertion order, gets tau = max(gauss(mean, sd), 0) as f32. */
void or_netb_set_synapse_tau_dist(or_netb *b, uint32_t sg, uint32_t tg, double mean, double sd,
                                  or_rng *r)
{
    if (sg >= b->ng || tg >= b->ng)
        return;
    uint8_t *in_s = (uint8_t *)calloc(b->n ? b->n : 1, 1), *in_t = (uint8_t *)calloc(b->n ? b->n : 1, 1);
    for (uint32_t k = 0; k < b->gr[sg].len; ++k) in_s[b->gr[sg].m[k]] = 1;
    for (uint32_t k = 0; k < b->gr[tg].len; ++k) in_t[b->gr[tg].m[k]] = 1;
    for (uint64_t j = 0; j < b->ns; ++j) {
        nb_syn *s = &b->sy[j];
        if (in_s[s->src] && in_t[s->tgt]) {
            double g = or_rng_gauss(r, mean, sd);
            s->tau_ms = (float)(g > 0.0 ? g : 0.0);
        }
    }
    free(in_s);
    free(in_t);
}

void or_netb_push_synapse(or_netb *b, uint32_t src, uint32_t tgt, float w, float delay_ms,
                          float tau_ms)
{
    if (b->ns == b->scap) {
        b->scap = b->scap ? 2 * b->scap : 1024;
        b->sy = (nb_syn *)realloc(b->sy, sizeof(nb_syn) * b->scap);
    }
    nb_syn *s = &b->sy[b->ns++];
    s->src = src;
    s->tgt = tgt;
    s->w = w;
    s->delay_ms = delay_ms;
    s->tau_ms = tau_ms;
}

/* Network::connect + conduction_delay_ms -- network.rs:596-613 */
uint64_t or_netb_connect(or_netb *b, uint32_t src, uint32_t tgt, float w)
{
    float d = pair_dist(&b->nr[src], &b->nr[tgt]);
    float delay = d / b->speed;
    if (delay < 1e-3f)
        delay = 1e-3f;
    or_netb_push_synapse(b, src, tgt, w, delay, 0.0f);
    return b->ns - 1;
}

static uint32_t new_group(or_netb *b, uint32_t cap)
{
    b->gr = (nb_group *)realloc(b->gr, sizeof(nb_group) * (b->ng + 1));
    b->gr[b->ng].m = (uint32_t *)malloc(sizeof(uint32_t) * (cap ? cap : 1));
    b->gr[b->ng].len = 0;
    return b->ng++;
}

/* builders::create_2d_grid -- builders.rs:26-68 */
uint32_t or_netb_grid2d(or_netb *b, uint32_t width, uint32_t height, float dist, float cx,
                        float cy, float cz)
{
    uint32_t g = new_group(b, width * height);
    for (uint32_t x = 0; x < width; ++x)
        for (uint32_t y = 0; y < height; ++y) {
            float py = cy + ((float)y - (float)height / 2.0f) * dist;
            float pz = cz + ((float)x - (float)width / 2.0f) * dist;
            uint32_t id = or_netb_add_neuron(b, cx, py, pz, 0);
            b->gr[g].m[b->gr[g].len++] = id;
        }
    return g;
}

/* builders::create_3d_grid -- builders.rs:88-131 */
uint32_t or_netb_grid3d(or_netb *b, uint32_t width, uint32_t height, uint32_t depth, float dist,
                        float cx, float cy, float cz)
{
    uint32_t g = new_group(b, width * height * depth);
    for (uint32_t z = 0; z < depth; ++z)
        for (uint32_t x = 0; x < width; ++x)
            for (uint32_t y = 0; y < height; ++y) {
                float px = cx + ((float)x - (float)width / 2.0f) * dist;
                float py = cy + ((float)y - (float)height / 2.0f) * dist;
                float pz = cz + ((float)z - (float)depth / 2.0f) * dist;
                uint32_t id = or_netb_add_neuron(b, px, py, pz, 0);
                b->gr[g].m[b->gr[g].len++] = id;
            }
    return g;
}

/* builders::create_connections -- builders.rs:150-181: one f64 draw per
 * ordered pair, p = scaling * exp(-(d/lambda)^2) in f64 */
void or_netb_create_connections(or_netb *b, uint32_t group, double scaling, double lambda,
                                float weight, or_rng *r)
{
    if (group >= b->ng)
        return;
    uint32_t len = b->gr[group].len;
    uint32_t *m = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
    memcpy(m, b->gr[group].m, sizeof(uint32_t) * len);
    for (uint32_t i = 0; i < len; ++i)
        for (uint32_t j = 0; j < len; ++j) {
            uint32_t n1 = m[i], n2 = m[j];
            if (n1 == n2)
                continue;
            double d = (double)pair_dist(&b->nr[n1], &b->nr[n2]);
            double q = d / lambda;
            if (or_rng_f64(r) < scaling * exp(-(q * q)))
                or_netb_connect(b, n1, n2, weight);
        }
    free(m);
}

/* builders::set_exc_inh -- builders.rs:262-277 */
void or_netb_set_exc_inh(or_netb *b, uint32_t group, double inh_fraction, or_rng *r)
{
    if (group >= b->ng)
        return;
    for (uint32_t k = 0; k < b->gr[group].len; ++k) {
        int kind = or_rng_f64(r) < inh_fraction ? 1 : 0;
        or_netb_set_kind(b, b->gr[group].m[k], kind);
    }
}

/* builders::set_exc_inh_exact -- builders.rs:280-296 (Fisher-Yates shuffle,
 * rng.rs:77-81) */
void or_netb_set_exc_inh_exact(or_netb *b, uint32_t group, uint64_t n_inh, or_rng *r)
{
    if (group >= b->ng)
        return;
    uint32_t len = b->gr[group].len;
    uint32_t *m = (uint32_t *)malloc(sizeof(uint32_t) * (len ? len : 1));
    memcpy(m, b->gr[group].m, sizeof(uint32_t) * len);
    for (uint32_t i = len ? len - 1 : 0; i >= 1; --i) {
        uint32_t j = (uint32_t)or_rng_below(r, (uint64_t)i + 1);
        uint32_t tmp = m[i];
        m[i] = m[j];
        m[j] = tmp;
    }
    for (uint32_t i = 0; i < len; ++i)
        or_netb_set_kind(b, m[i], i < n_inh ? 1 : 0);
    free(m);
}

/* bench::bench_net -- sim/src/bench.rs:10-53 */
void or_netb_bench_net(or_netb *b, uint32_t n, uint32_t k, uint64_t seed, float w_scale,
                       float w_inh)
{
    or_rng r;
    or_rng_init(&r, seed);
    b->speed = 1000.0f;
    uint64_t n_exc = (uint64_t)n * 4 / 5;
    for (uint32_t i = 0; i < n; ++i) {
        float x = (float)or_rng_uniform(&r, -500.0, 500.0);
        float y = (float)or_rng_uniform(&r, -500.0, 500.0);
        float z = (float)or_rng_uniform(&r, -500.0, 500.0);
        or_netb_add_neuron(b, x, y, z, i < n_exc ? 0 : 1);
    }
    for (uint32_t src = 0; src < n; ++src) {
        int exc = src < n_exc;
        float w = exc ? 6.0f * w_scale : w_inh;
        for (uint32_t q = 0; q < k; ++q) {
            uint32_t tgt = (uint32_t)or_rng_below(&r, n);
            float delay = exc ? (float)or_rng_uniform(&r, 1.0, 21.0) : 1.0f;
            or_netb_push_synapse(b, src, tgt, w, delay, 0.0f);
        }
    }
}

/* Network::add_input -- network.rs:719-731 */
uint32_t or_netb_add_input(or_netb *b, uint32_t n_targets, const uint32_t *tgt, const float *w)
{
    b->in = (nb_input *)realloc(b->in, sizeof(nb_input) * (b->ni + 1));
    nb_input *x = &b->in[b->ni];
    x->tgt = (uint32_t *)malloc(sizeof(uint32_t) * (n_targets ? n_targets : 1));
    x->w = (float *)malloc(sizeof(float) * (n_targets ? n_targets : 1));
    memcpy(x->tgt, tgt, sizeof(uint32_t) * n_targets);
    memcpy(x->w, w, sizeof(float) * n_targets);
    x->len = n_targets;
    return b->ni++;
}

/* Network::synapse_csr -- network.rs:654-692: stable by src, Dale sign from
 * the source kind, delay steps = round(delay/dt) min 1, error past u16. */
int or_netb_csr(const or_netb *b, float dt_ms, uint32_t *src_ptr, uint32_t *tgt, float *w,
                uint16_t *delay, float *tau, char *err, int errlen)
{
    uint32_t n = b->n;
    memset(src_ptr, 0, sizeof(uint32_t) * (n + 1));
    for (uint64_t j = 0; j < b->ns; ++j)
        src_ptr[b->sy[j].src + 1]++;
    for (uint32_t i = 0; i < n; ++i)
        src_ptr[i + 1] += src_ptr[i];
    uint32_t *fill = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    for (uint64_t j = 0; j < b->ns; ++j) {
        const nb_syn *s = &b->sy[j];
        float steps = roundf(s->delay_ms / dt_ms);
        if (steps < 1.0f)
            steps = 1.0f;
        if (!isfinite(steps) || steps > 65535.0f) {
            snprintf(err, errlen,
                     "synapse %u -> %u: delay %g ms = %g steps at dt %g exceeds the 65535-step ring",
                     s->src, s->tgt, (double)s->delay_ms, (double)steps, (double)dt_ms);
            free(fill);
            return -1;
        }
        uint64_t o = src_ptr[s->src] + fill[s->src]++;
        tgt[o] = s->tgt;
        int inh = b->nr[s->src].kind == 1;
        w[o] = inh ? -fabsf(s->w) : fabsf(s->w);
        delay[o] = (uint16_t)steps;
        if (tau)
            tau[o] = s->tau_ms;
    }
    free(fill);
    return 0;
}

/* NeuronParams::from_network -- sim/src/state.rs:98-153 (model table in
 * first-use order, rows padded to the widest model) + model init
 * (izh.rs:300-303, adex.rs:149-152). */
void or_netb_params(const or_netb *b, uint32_t *n_models, uint8_t *model_ids, uint8_t *model,
                    uint32_t *n_params, uint32_t *n_state, float *p, float *x0)
{
    uint32_t nm = 0, np = 0, nx = 0;
    for (uint32_t i = 0; i < b->n; ++i) {
        const nb_neuron *r = &b->nr[i];
        uint32_t k;
        for (k = 0; k < nm; ++k)
            if (model_ids[k] == r->model)
                break;
        if (k == nm)
            model_ids[nm++] = (uint8_t)r->model;
        model[i] = (uint8_t)k;
        uint32_t w = r->model == 1 ? 10 : (r->model == 2 ? 5 : 4);
        if (w > np)
            np = w;
        nx = 2;
    }
    for (uint32_t i = 0; i < b->n; ++i) {
        const nb_neuron *r = &b->nr[i];
        for (uint32_t q = 0; q < np; ++q)
            p[(uint64_t)i * np + q] = q < (uint32_t)r->np ? r->p[q] : 0.0f;
        float *x = x0 + (uint64_t)i * nx;
        if (r->model == 1) {
            x[0] = r->p[2];
            x[1] = 0.0f;
        } else if (r->model == 2) {
            /* the LIF test model's init (sim/tests/lif.rs:95-97): v = e */
            x[0] = r->p[1];
            x[1] = 0.0f;
        } else {
            x[0] = -65.0f;
            x[1] = r->p[1] * x[0];
        }
    }
    *n_models = nm;
    *n_params = np;
    *n_state = nx;
}

uint32_t or_netb_n_inputs(const or_netb *b) { return b->ni; }

uint64_t or_netb_n_stim(const or_netb *b)
{
    uint64_t c = 0;
    for (uint32_t k = 0; k < b->ni; ++k)
        for (uint32_t q = 0; q < b->in[k].len; ++q)
            if (b->in[k].tgt[q] < b->n)
                ++c;
    return c;
}

/* StimMap::from_network -- sim/src/stim.rs:34-57 */
void or_netb_stim_map(const or_netb *b, uint32_t *stim_ptr, uint32_t *stim_input, float *stim_w)
{
    uint32_t n = b->n;
    memset(stim_ptr, 0, sizeof(uint32_t) * (n + 1));
    for (uint32_t k = 0; k < b->ni; ++k)
        for (uint32_t q = 0; q < b->in[k].len; ++q)
            if (b->in[k].tgt[q] < n)
                stim_ptr[b->in[k].tgt[q] + 1]++;
    for (uint32_t i = 0; i < n; ++i)
        stim_ptr[i + 1] += stim_ptr[i];
    uint32_t *fill = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    for (uint32_t k = 0; k < b->ni; ++k)
        for (uint32_t q = 0; q < b->in[k].len; ++q) {
            uint32_t t = b->in[k].tgt[q];
            if (t >= n)
                continue;
            uint32_t o = stim_ptr[t] + fill[t]++;
            stim_input[o] = k;
            stim_w[o] = b->in[k].w[q];
        }
    free(fill);
}

void or_netb_positions(const or_netb *b, float *xyz)
{
    for (uint32_t i = 0; i < b->n; ++i)
        memcpy(xyz + 3 * (uint64_t)i, b->nr[i].pos, sizeof(float) * 3);
}

void or_netb_kinds(const or_netb *b, uint8_t *kind)
{
    for (uint32_t i = 0; i < b->n; ++i)
        kind[i] = (uint8_t)b->nr[i].kind;
}
