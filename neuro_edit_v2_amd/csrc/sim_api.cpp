// sim_api.cpp -- nh_sim: the single-network backend behind the C ABI.
//
// Mirrors the reference Backend trait (sim/src/lib.rs:119-184) as
// implemented by CudaSim (sim/src/cuda.rs): the stimulus program, the
// external-current host mirror, watches and readbacks are host-side; the
// step itself runs on the arena's stream.
//
// Readouts that the reference keeps as per-step device state are derived
// from per-neuron last-spike steps instead, with the same arithmetic:
//  * trace (cpu.rs:252-258): 1 on spike, times decay_t every later step --
//    rebuilt on read by the same repeated f32 multiply;
//  * grid traces (cpu.rs:293-303, 391-407): last spike per texel since the
//    watch was set, exp on the host with the f32 libm exp the CPU backend
//    uses.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "arena.h"

namespace nh {
const char *last_error_text();
}

using nh::Arena;
using nh::set_error;

struct nh_sim {
    Arena ar;
    nh_config cfg{};
    uint32_t n = 0, n_inputs = 0;
    // stimulus program (sim/src/stim.rs:9-21)
    std::vector<float> constants;
    std::vector<std::vector<float>> series;
    std::vector<uint64_t> play_start;
    std::vector<uint8_t> looped;
    std::vector<float> ext_host;
    // watches
    std::vector<uint32_t> v_watch;
    uint64_t vrec_valid = 0;
    std::vector<std::vector<uint32_t>> grid_groups;
    uint64_t grid_epoch = 0;
    // trace power table: pw[k] = decay_t multiplied k times into 1.0
    std::vector<float> trace_pw;
    float decay_t = 0.f;
};

#define TRY(expr)            \
    do {                     \
        int rc_ = (expr);    \
        if (rc_) return rc_; \
    } while (0)

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(NH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHECK_SIM(s) \
    if (!(s)) return set_error(NH_ERR_ARG, "null sim")

// Stimulus::value -- sim/src/stim.rs:102-115
static float stim_value(const nh_sim *s, uint32_t k, uint64_t t) {
    if (k < s->series.size() && !s->series[k].empty()) {
        const auto &ser = s->series[k];
        const uint64_t start = s->play_start[k];
        if (s->looped[k]) {
            const int64_t len = (int64_t)ser.size();
            int64_t m = ((int64_t)t - (int64_t)start) % len;
            if (m < 0) m += len;
            return ser[(size_t)m];
        }
        const uint64_t idx = t - start;
        return idx < ser.size() ? ser[idx] : 0.0f;
    }
    return k < s->constants.size() ? s->constants[k] : 0.0f;
}

// Stimulus::pack -- sim/src/stim.rs:119-127
static void pack_stim(void *ctx, uint64_t t0, uint32_t k, float *out) {
    const nh_sim *s = (const nh_sim *)ctx;
    for (uint32_t j = 0; j < k; ++j)
        for (uint32_t q = 0; q < s->n_inputs; ++q) out[(size_t)j * s->n_inputs + q] = stim_value(s, q, t0 + j);
}

extern "C" {

const char *nh_last_error(void) { return nh::last_error_text(); }

void nh_config_default(nh_config *c) {
    std::memset(c, 0, sizeof(*c));
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
    c->raster_steps = 0;
    c->exec_mode = 0;
    c->step_shape = -1;
    c->graph_steps = 0;
    c->block = 0;
    c->lds_bins = 0;
}

int nh_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 0;
    return 1;
}

int nh_check_net(const nh_net_desc *net, const nh_config *cfg) {
    if (!net || !cfg) return set_error(NH_ERR_ARG, "null argument");
    if (!(cfg->dt_ms >= 0.001f && cfg->dt_ms <= 100.0f))
        return set_error(NH_ERR_ARG, "dt_ms outside 0.001..100 (sim/src/lib.rs:35-39)");
    return nh::check_desc(*net, *cfg, nullptr);
}

int nh_create(const nh_net_desc *net, const nh_config *cfg, nh_sim **out) {
    return nh::guarded([&]() -> int {
        if (!net || !cfg || !out) return set_error(NH_ERR_ARG, "null argument");
        *out = nullptr;
        TRY(nh_check_net(net, cfg));
        if (!nh_available()) return set_error(NH_ERR_NO_DEVICE, "no HIP device");
        std::unique_ptr<nh_sim> s(new (std::nothrow) nh_sim());
        if (!s) return set_error(NH_ERR_CAPACITY, "out of host memory");
        s->cfg = *cfg;
        s->n = net->n;
        s->n_inputs = net->n_inputs;
        nh::Tuning tune;
        tune.fused = cfg->step_shape;
        tune.use_graph = cfg->exec_mode == 0 ? -1 : (cfg->exec_mode == 2 ? 1 : 0);
        tune.graph_steps = cfg->graph_steps ? cfg->graph_steps : 64;
        tune.block = cfg->block;
        tune.lds_bins = cfg->lds_bins;
        TRY(s->ar.build(1, net, cfg, cfg->raster_steps, tune));
        s->constants.assign(s->n_inputs, 0.0f);
        s->series.assign(s->n_inputs, {});
        s->play_start.assign(s->n_inputs, 0);
        s->looped.assign(s->n_inputs, 0);
        s->ext_host.assign(s->n, 0.0f);
        s->ar.set_stim_packer(pack_stim, s.get());
        const float rt = cfg->readout_tau_ms > 1e-6f ? cfg->readout_tau_ms : 1e-6f;
        s->decay_t = std::exp(-cfg->dt_ms / rt);  // cpu.rs:204
        *out = s.release();
        return NH_OK;
    });
}

void nh_destroy(nh_sim *s) { delete s; }

int nh_run_steps(nh_sim *s, uint64_t n_steps) {
    return nh::guarded([&]() -> int {
        CHECK_SIM(s);
        TRY(s->ar.run(n_steps));
        if (!s->v_watch.empty()) s->vrec_valid += n_steps;
        return NH_OK;
    });
}

int nh_run_ms(nh_sim *s, float ms) {
    CHECK_SIM(s);
    // cpu.rs:409-412: steps = round(ms / dt)
    const float steps = std::round(ms / s->cfg.dt_ms);
    if (!(steps >= 0.0f)) return set_error(NH_ERR_ARG, "negative run length");
    return nh_run_steps(s, (uint64_t)steps);
}

int nh_reset(nh_sim *s) {
    CHECK_SIM(s);
    TRY(s->ar.reset());
    s->vrec_valid = 0;
    s->grid_epoch = 0;
    return NH_OK;
}

int nh_sync(nh_sim *s) {
    CHECK_SIM(s);
    return s->ar.sync();
}

uint32_t nh_n(const nh_sim *s) { return s ? s->n : 0; }
uint64_t nh_step_count(const nh_sim *s) { return s ? s->ar.t : 0; }
double nh_time_ms(const nh_sim *s) { return s ? (double)s->ar.t * (double)s->cfg.dt_ms : 0.0; }
uint64_t nh_dev_mem(const nh_sim *s) { return s ? s->ar.dev_bytes : 0; }

int nh_last_run_ms(nh_sim *s, double *ms) {
    CHECK_SIM(s);
    TRY(s->ar.sync_timed());
    *ms = s->ar.last_run_ms;
    return NH_OK;
}

int nh_last_run_tail(nh_sim *s, double *ms, uint64_t *launches) {
    CHECK_SIM(s);
    TRY(s->ar.sync_timed());
    *ms = s->ar.last_tail_ms;
    *launches = s->ar.last_tail_launches;
    return NH_OK;
}

int nh_profile_steps(nh_sim *s, uint64_t n_steps, double *main_ms, uint64_t *main_n, double *aux_ms,
                     uint64_t *aux_n) {
    CHECK_SIM(s);
    TRY(s->ar.profile(n_steps, *main_ms, *main_n, *aux_ms, *aux_n));
    if (!s->v_watch.empty()) s->vrec_valid += n_steps;
    return NH_OK;
}

int nh_test_pair_wait(nh_sim *s, uint32_t spin_limit, uint32_t mute_wave0, uint64_t *fallbacks) {
    CHECK_SIM(s);
    TRY(s->ar.sync());
    if (fallbacks) *fallbacks = s->ar.pair_fallbacks;
    s->ar.pair_spins = spin_limit;
    s->ar.pair_mute = mute_wave0 ? 1u : 0u;
    s->ar.graph_dirty = true;  // both ride in the captured kernel arguments
    return NH_OK;
}

int nh_launch_shape(nh_sim *s, int32_t *shape) {
    CHECK_SIM(s);
    if (!shape) return set_error(NH_ERR_ARG, "null output");
    *shape = s->ar.run_usable() ? 3 : (s->ar.pair_usable() ? 2 : (s->ar.fused ? 1 : 0));
    return NH_OK;
}

int nh_autotune(nh_sim *s, uint32_t trial_steps, char *report, uint32_t report_cap) {
    CHECK_SIM(s);
    std::string rep;
    TRY(s->ar.autotune(trial_steps ? trial_steps : 256, rep));
    s->vrec_valid = 0;
    s->grid_epoch = 0;
    if (report && report_cap) {
        std::strncpy(report, rep.c_str(), report_cap - 1);
        report[report_cap - 1] = 0;
    }
    return NH_OK;
}

int nh_probe_copy(uint64_t bytes, uint32_t reps, uint32_t lane_bytes, double *ms) {
    if (!ms || bytes < 16 || reps == 0 || (lane_bytes != 4 && lane_bytes != 16))
        return set_error(NH_ERR_ARG, "bad probe arguments");
    if (!nh_available()) return set_error(NH_ERR_NO_DEVICE, "no HIP device");
    bytes &= ~uint64_t(15);
    nh::DBuf a, b;
    HIPCHK(a.alloc(bytes));
    HIPCHK(b.alloc(bytes));
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    nh::launch_copy(a.p, b.p, bytes, lane_bytes, st);  // warm
    HIPCHK(hipEventRecord(e0, st));
    for (uint32_t r = 0; r < reps; ++r) nh::launch_copy(a.p, b.p, bytes, lane_bytes, st);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipStreamSynchronize(st));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    return NH_OK;
}

int nh_probe_launch(uint32_t blocks, uint32_t launches, double *us) {
    if (!us || blocks == 0 || launches == 0 || launches > 100000)
        return set_error(NH_ERR_ARG, "bad probe arguments");
    if (!nh_available()) return set_error(NH_ERR_NO_DEVICE, "no HIP device");
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipGraph_t g;
    hipGraphExec_t ge;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (uint32_t k = 0; k < launches; ++k) nh::launch_empty(blocks, st);
    HIPCHK(hipStreamEndCapture(st, &g));
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipGraphLaunch(ge, st));  // warm
    HIPCHK(hipEventRecord(e0, st));
    const int reps = 5;
    for (int r = 0; r < reps; ++r) HIPCHK(hipGraphLaunch(ge, st));
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipStreamSynchronize(st));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *us = (double)t * 1e3 / ((double)reps * launches);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(st);
    return NH_OK;
}

static int read_last(nh_sim *s, std::vector<uint32_t> &last) {
    return s->ar.read_u32(s->ar.d_last, last, s->n);
}

int nh_read_trace(nh_sim *s, float *out, uint32_t n) {
    CHECK_SIM(s);
    if (n < s->n) return set_error(NH_ERR_ARG, "output too small");
    std::vector<uint32_t> last;
    TRY(read_last(s, last));
    const uint64_t t = s->ar.t;
    uint64_t maxgap = 0;
    for (uint32_t i = 0; i < s->n; ++i)
        if (last[i] != nh::kNoSpike) maxgap = std::max<uint64_t>(maxgap, t - 1 - last[i]);
    auto &pw = s->trace_pw;
    if (pw.empty()) pw.push_back(1.0f);
    while (pw.size() <= maxgap && pw.back() != 0.0f) pw.push_back(pw.back() * s->decay_t);
    for (uint32_t i = 0; i < s->n; ++i) {
        if (last[i] == nh::kNoSpike) {
            out[i] = 0.0f;
            continue;
        }
        const uint64_t g = t - 1 - last[i];
        out[i] = g < pw.size() ? pw[g] : 0.0f;
    }
    return NH_OK;
}

int nh_read_state(nh_sim *s, uint32_t slot, float *out, uint32_t n) {
    CHECK_SIM(s);
    if (slot >= 2) return set_error(NH_ERR_ARG, "state slot out of range");
    if (n < s->n) return set_error(NH_ERR_ARG, "output too small");
    std::vector<float2> x;
    TRY(s->ar.read_x(x));
    for (uint32_t i = 0; i < s->n; ++i) out[i] = slot == 0 ? x[i].x : x[i].y;
    return NH_OK;
}

int nh_read_syn_current(nh_sim *s, float *out, uint32_t n) {
    CHECK_SIM(s);
    if (n < s->n) return set_error(NH_ERR_ARG, "output too small");
    std::vector<float> v;
    TRY(s->ar.read_f32(s->ar.d_s, v, s->n));
    std::memcpy(out, v.data(), (size_t)s->n * 4);
    return NH_OK;
}

int nh_spikes_total(nh_sim *s, uint64_t *out) {
    CHECK_SIM(s);
    return s->ar.spikes_total(*out);
}

int nh_events_total(nh_sim *s, uint64_t *out) {
    CHECK_SIM(s);
    return s->ar.events_total(*out);
}

int nh_mark_counters(nh_sim *s) {
    return nh::guarded([&]() -> int {
        CHECK_SIM(s);
        return s->ar.mark_counters();
    });
}

int nh_counts_since_mark(nh_sim *s, uint64_t *spikes, uint64_t *events) {
    return nh::guarded([&]() -> int {
        CHECK_SIM(s);
        if (!spikes || !events) return set_error(NH_ERR_ARG, "null argument");
        return s->ar.counts_since_mark(*spikes, *events);
    });
}

int nh_mean_rate_hz(nh_sim *s, double *out) {
    CHECK_SIM(s);
    *out = 0.0;
    if (s->ar.t == 0) return NH_OK;
    uint64_t tot = 0;
    TRY(s->ar.spikes_total(tot));
    *out = (double)tot / (double)std::max<uint32_t>(s->n, 1) /
           ((double)s->ar.t * (double)s->cfg.dt_ms / 1000.0);
    return NH_OK;
}

int nh_read_pn_total(nh_sim *s, uint64_t *out, uint32_t n) {
    CHECK_SIM(s);
    if (n < s->n) return set_error(NH_ERR_ARG, "output too small");
    std::vector<uint32_t> v;
    TRY(s->ar.read_u32(s->ar.d_pn, v, s->n));
    for (uint32_t i = 0; i < s->n; ++i) out[i] = v[i];
    return NH_OK;
}

// CpuSim::spike_row -- cpu.rs:429-440
int nh_spike_row(nh_sim *s, uint64_t back, uint64_t *words, uint32_t n_words) {
    CHECK_SIM(s);
    const uint32_t W = (s->n + 63) / 64;
    if (n_words < W) return set_error(NH_ERR_ARG, "output too small");
    const uint64_t t = s->ar.t;
    const uint32_t R = s->ar.bits_rows;
    if (!s->cfg.record_spikes || t == 0 || back >= t || back >= R || R == 0)
        return set_error(NH_ERR_NONE, "no recorded row");
    TRY(s->ar.sync());
    const uint64_t row = (t - 1 - back) % R;
    if (W)
        HIPCHK(hipMemcpy(words, s->ar.d_bits.as<uint64_t>() + row * s->ar.words, (size_t)W * 8,
                         hipMemcpyDeviceToHost));
    // bits past n are padding lanes that never spike
    return NH_OK;
}

int nh_last_spikes(nh_sim *s, uint32_t *ids, uint32_t cap, uint32_t *n_out) {
    CHECK_SIM(s);
    *n_out = 0;
    const uint32_t W = (s->n + 63) / 64;
    std::vector<uint64_t> row(std::max<uint32_t>(W, 1));
    int rc = nh_spike_row(s, 0, row.data(), W);
    if (rc == NH_ERR_NONE) return NH_OK;
    if (rc) return rc;
    uint32_t c = 0;
    for (uint32_t wi = 0; wi < W; ++wi) {
        uint64_t w = row[wi];
        while (w) {
            const uint32_t b = (uint32_t)__builtin_ctzll(w);
            w &= w - 1;
            const uint32_t i = wi * 64 + b;
            if (i < s->n) {
                if (c >= cap) return set_error(NH_ERR_CAPACITY, "ids buffer too small");
                ids[c++] = i;
            }
        }
    }
    *n_out = c;
    return NH_OK;
}

// CpuSim::spike_counts -- cpu.rs:452-486; CudaSim::spike_counts
// (cuda.rs:725-763) copies the whole record ring and counts on the host,
// here the window is counted on the device and only n counts are copied.
int nh_spike_counts(nh_sim *s, uint64_t rows, uint32_t *out, uint32_t n) {
    return nh::guarded([&]() -> int {
        CHECK_SIM(s);
        if (!out) return set_error(NH_ERR_ARG, "null output");
        if (n < s->n) return set_error(NH_ERR_ARG, "output too small");
        std::memset(out, 0, (size_t)s->n * 4);
        const uint64_t t = s->ar.t;
        if (!s->cfg.record_spikes || t == 0 || s->n == 0 || rows == 0) return NH_OK;
        if (rows >= t) return s->ar.read_pn(0, s->n, out);
        const uint32_t R = s->ar.bits_rows;
        if (rows > R)
            return set_error(NH_ERR_WINDOW, "spike_counts window " + std::to_string(rows) +
                                                " exceeds the recorded " + std::to_string(R) + " steps");
        if (!s->ar.d_counts.p) {
            const hipError_t e = s->ar.d_counts.alloc((size_t)s->ar.n_total * 4);
            if (e != hipSuccess) return set_error(NH_ERR_HIP, std::string("count buffer: ") + hipGetErrorString(e));
            s->ar.dev_bytes += s->ar.d_counts.bytes;
        }
        return s->ar.count_rows(0, s->n, rows, s->ar.d_counts.as<uint32_t>(), out);
    });
}

// nh_spike_counts without device allocation: the window is counted into the
// library's file-scope scratch (kScratchWords counts per device, shared by
// every call on that device: calls that overlap on one GPU race).  Every
// count the caller receives is written by the counting kernel of this call,
// so nothing from an earlier call needs clearing.
int nh_spike_counts_noalloc(nh_sim *s, uint64_t rows, uint32_t *out, uint32_t n) {
    return nh::guarded([&]() -> int {
        CHECK_SIM(s);
        if (!out) return set_error(NH_ERR_ARG, "null output");
        if (n < s->n) return set_error(NH_ERR_ARG, "output too small");
        if (s->n > nh::kScratchWords)
            return set_error(NH_ERR_CAPACITY, "spike_counts_noalloc: " + std::to_string(s->n) +
                                                  " neurons exceed the scratch of " +
                                                  std::to_string(nh::kScratchWords) + " counts");
        std::memset(out, 0, (size_t)s->n * 4);
        const uint64_t t = s->ar.t;
        if (!s->cfg.record_spikes || t == 0 || s->n == 0 || rows == 0) return NH_OK;
        if (rows >= t) return s->ar.read_pn(0, s->n, out);
        const uint32_t R = s->ar.bits_rows;
        if (rows > R)
            return set_error(NH_ERR_WINDOW, "spike_counts window " + std::to_string(rows) +
                                                " exceeds the recorded " + std::to_string(R) + " steps");
        uint32_t *scr = nullptr;
        const hipError_t e = nh::scratch_on_device(s->ar.device, &scr);
        if (e != hipSuccess) return set_error(NH_ERR_HIP, std::string("count scratch: ") + hipGetErrorString(e));
        return s->ar.count_rows(0, s->n, rows, scr, out);
    });
}

int nh_set_input_constant(nh_sim *s, uint32_t input, float value) {
    CHECK_SIM(s);
    if (input >= s->n_inputs) return set_error(NH_ERR_ARG, "input out of range");
    s->constants[input] = value;
    return NH_OK;
}

int nh_input_values(nh_sim *s, float *out, uint32_t n_inputs) {
    CHECK_SIM(s);
    if (n_inputs < s->n_inputs) return set_error(NH_ERR_ARG, "output too small");
    std::copy(s->constants.begin(), s->constants.end(), out);
    return NH_OK;
}

static int play(nh_sim *s, uint32_t input, const float *series, uint64_t len, bool looped) {
    CHECK_SIM(s);
    if (input >= s->n_inputs) return set_error(NH_ERR_ARG, "input out of range");
    s->series[input].assign(series, series + len);
    s->play_start[input] = s->ar.t;
    s->looped[input] = looped ? 1 : 0;
    return NH_OK;
}

int nh_play_input(nh_sim *s, uint32_t input, const float *series, uint64_t len) {
    return play(s, input, series, len, false);
}

int nh_loop_input(nh_sim *s, uint32_t input, const float *series, uint64_t len) {
    return play(s, input, series, len, true);
}

static int push_ext(nh_sim *s) {
    bool any = false;
    for (float v : s->ext_host)
        if (v != 0.0f) {
            any = true;
            break;
        }
    // ext is read only while some entry is nonzero; s + 0.0f == s + ext[i]
    // for an all-zero ext, so dropping the read changes no bit.
    if (any != s->ar.has_ext) {
        s->ar.has_ext = any;
        s->ar.graph_dirty = true;
    }
    HIPCHK(hipMemcpyAsync(s->ar.d_ext.p, s->ext_host.data(), (size_t)s->n * 4, hipMemcpyHostToDevice,
                          s->ar.stream));
    HIPCHK(hipStreamSynchronize(s->ar.stream));
    return NH_OK;
}

int nh_set_external(nh_sim *s, uint32_t neuron, float amps) {
    CHECK_SIM(s);
    if (neuron >= s->n) return NH_OK;  // lib.rs: out-of-range is ignored (cpu.rs:590-594)
    s->ext_host[neuron] = amps;
    return push_ext(s);
}

int nh_set_external_all(nh_sim *s, const float *ext, uint32_t n) {
    CHECK_SIM(s);
    const uint32_t m = std::min(n, s->n);
    std::copy(ext, ext + m, s->ext_host.begin());
    return push_ext(s);
}

int nh_read_external(nh_sim *s, float *out, uint32_t n) {
    CHECK_SIM(s);
    if (n < s->n) return set_error(NH_ERR_ARG, "output too small");
    std::copy(s->ext_host.begin(), s->ext_host.end(), out);
    return NH_OK;
}

int nh_set_noise(nh_sim *s, float rate_hz, float kick_mv) {
    CHECK_SIM(s);
    TRY(s->ar.sync());
    s->cfg.noise_rate_hz = rate_hz;
    s->cfg.noise_kick_mv = kick_mv;
    Arena::fill_noise(s->ar.recs[0], s->cfg);
    s->ar.graph_dirty = true;  // the record rides in the captured kernel arguments
    return s->ar.upload_recs();
}

int nh_set_noise_sd(nh_sim *s, float sd) {
    CHECK_SIM(s);
    TRY(s->ar.sync());
    s->cfg.noise_sd = sd;
    Arena::fill_noise(s->ar.recs[0], s->cfg);
    s->ar.graph_dirty = true;
    return s->ar.upload_recs();
}

int nh_set_v_watch(nh_sim *s, const uint32_t *ids, uint32_t n_ids) {
    CHECK_SIM(s);
    std::vector<uint32_t> v(ids, ids + n_ids);
    if (v == s->v_watch) return NH_OK;  // no-op when unchanged (cpu.rs:330-333)
    for (uint32_t id : v)
        if (id >= s->n) return set_error(NH_ERR_ARG, "watched neuron out of range");
    TRY(s->ar.sync());
    s->v_watch = v;
    s->vrec_valid = 0;
    auto &ar = s->ar;
    ar.n_watch = n_ids;
    ar.d_watch.release();
    ar.d_xrec.release();
    if (n_ids) {
        HIPCHK(ar.d_watch.alloc((size_t)n_ids * 4));
        HIPCHK(hipMemcpy(ar.d_watch.p, ids, (size_t)n_ids * 4, hipMemcpyHostToDevice));
        HIPCHK(ar.d_xrec.alloc((size_t)ar.xrec_depth * n_ids * 8));
    }
    ar.graph_dirty = true;
    return NH_OK;
}

// drain: out[slot][watch][cap_samples] (sim/src/cuda.rs:537-570 semantics)
int nh_drain_watch(nh_sim *s, float *out, uint32_t cap_samples, uint32_t *n_samples) {
    CHECK_SIM(s);
    *n_samples = 0;
    const uint32_t W = (uint32_t)s->v_watch.size();
    if (W == 0 || s->vrec_valid == 0) return NH_OK;
    auto &ar = s->ar;
    TRY(ar.sync());
    const uint64_t k = std::min<uint64_t>(s->vrec_valid, ar.xrec_depth);
    if (k > cap_samples) return set_error(NH_ERR_CAPACITY, "sample buffer too small");
    std::vector<float2> ring((size_t)ar.xrec_depth * W);
    HIPCHK(hipMemcpy(ring.data(), ar.d_xrec.p, ring.size() * 8, hipMemcpyDeviceToHost));
    const uint64_t t = ar.t;
    for (uint64_t r = 0; r < k; ++r) {
        const uint64_t step = t - k + r;
        const uint64_t row = step % ar.xrec_depth;
        for (uint32_t c = 0; c < W; ++c) {
            const float2 v = ring[row * W + c];
            out[((size_t)0 * W + c) * cap_samples + r] = v.x;
            out[((size_t)1 * W + c) * cap_samples + r] = v.y;
        }
    }
    *n_samples = (uint32_t)k;
    s->vrec_valid = 0;
    return NH_OK;
}

int nh_set_grid_watches(nh_sim *s, uint32_t n_groups, const uint32_t *sizes, const uint32_t *members) {
    CHECK_SIM(s);
    std::vector<std::vector<uint32_t>> g(n_groups);
    uint64_t o = 0;
    for (uint32_t k = 0; k < n_groups; ++k) {
        g[k].assign(members + o, members + o + sizes[k]);
        o += sizes[k];
    }
    if (g == s->grid_groups) return NH_OK;  // no-op when unchanged (cpu.rs:361-370)
    s->grid_groups = std::move(g);
    s->grid_epoch = s->ar.t;
    return NH_OK;
}

int nh_grid_traces(nh_sim *s, float tau_ms, float *out, uint64_t cap) {
    CHECK_SIM(s);
    uint64_t total = 0;
    for (auto &g : s->grid_groups) total += g.size();
    if (cap < total) return set_error(NH_ERR_ARG, "output too small");
    std::vector<uint32_t> last;
    TRY(read_last(s, last));
    const float kk = s->cfg.dt_ms / (tau_ms > 1e-6f ? tau_ms : 1e-6f);
    const uint64_t t = s->ar.t;
    uint64_t o = 0;
    std::vector<int32_t> map(s->n);
    for (auto &g : s->grid_groups) {
        std::fill(map.begin(), map.end(), -1);
        for (uint32_t k = 0; k < g.size(); ++k)
            if (g[k] < s->n) map[g[k]] = (int32_t)k;
        std::vector<uint32_t> ls(g.size(), 0xFFFFFFFFu);
        for (uint32_t k = 0; k < g.size(); ++k) {
            const uint32_t m = g[k];
            if (m >= s->n || map[m] != (int32_t)k) continue;
            const uint32_t l = last[m];
            if (l == nh::kNoSpike || (uint64_t)l < s->grid_epoch) continue;
            ls[k] = l;
        }
        for (uint32_t k = 0; k < g.size(); ++k) {
            if (ls[k] == 0xFFFFFFFFu) {
                out[o++] = 0.0f;
            } else {
                const uint64_t gap = t > ls[k] ? t - ls[k] : 0;
                out[o++] = std::exp(-((float)gap) * kk);
            }
        }
    }
    return NH_OK;
}

int nh_read_spikes(nh_sim *s, uint64_t *step0, uint32_t *counts, uint64_t max_steps, uint32_t *ids,
                   uint64_t max_ids, uint64_t *n_steps, uint64_t *n_ids) {
    return nh::guarded([&]() -> int {
        CHECK_SIM(s);
        auto &ar = s->ar;
        if (!ar.rast_steps) return set_error(NH_ERR_UNSUPPORTED, "raster stream off (raster_steps = 0)");
        TRY(ar.drain_raster());
        *n_steps = ar.h_rast_cnt.size();
        *n_ids = ar.h_rast_ids.size();
        *step0 = ar.h_rast_step0;
        if (*n_steps > max_steps || *n_ids > max_ids)
            return set_error(NH_ERR_CAPACITY, "raster output arrays too small");
        std::copy(ar.h_rast_cnt.begin(), ar.h_rast_cnt.end(), counts);
        std::copy(ar.h_rast_ids.begin(), ar.h_rast_ids.end(), ids);
        ar.h_rast_cnt.clear();
        ar.h_rast_ids.clear();
        ar.h_rast_step0 = ar.t;
        return NH_OK;
    });
}

}  // extern "C"
