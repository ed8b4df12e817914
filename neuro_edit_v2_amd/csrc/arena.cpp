// arena.cpp -- build, launch sequence and readbacks of a packed arena.
//
// Construction follows CudaSim::with_params (sim/src/cuda.rs:130-239): the
// CSR, ring depth (state.rs:221-224), stimulus map and parameter rows are
// flattened once and uploaded; afterwards the per-step work is enqueued on
// the arena's stream without host synchronisation (cuda.rs:584-604).
#include "arena.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>

namespace nh {

namespace {
thread_local std::string g_err;
}

int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
const char *last_error_text() { return g_err.c_str(); }

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(NH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

Arena::~Arena() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int i = 0; i < 2; ++i) {
        if (h_stage[i]) (void)hipHostFree(h_stage[i]);
        if (stage_ev[i]) (void)hipEventDestroy(stage_ev[i]);
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_tail) (void)hipEventDestroy(ev_tail);
    if (stream) (void)hipStreamDestroy(stream);
}

int Arena::alloc(DBuf &b, size_t bytes) {
    HIPCHK(b.alloc(bytes));
    dev_bytes += b.bytes;
    return NH_OK;
}

template <class T> int Arena::upload(DBuf &b, const std::vector<T> &v) {
    int rc = alloc(b, v.size() * sizeof(T));
    if (rc) return rc;
    if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return NH_OK;
}

#define TRY(expr)                \
    do {                         \
        int rc_ = (expr);        \
        if (rc_) return rc_;     \
    } while (0)

static uint32_t align64(uint64_t v) { return (uint32_t)((v + 63) & ~uint64_t(63)); }

// sim/src/kernels.rs:61 -- (int)lrintf(w * WSCALE), round half to even.
static int32_t quantize(float w, float scale) { return (int32_t)std::lrintf(w * scale); }

// Host-side checks of one network and its configuration: everything the
// device path assumes about the arrays (NeuronParams::check,
// sim/src/state.rs:197-217; the per-synapse and per-input checks of
// Network::validate, net-format/src/network.rs:1074-1150, on the flat
// layout), made before any device call.  *max_delay gets the largest delay.
int check_desc(const nh_net_desc &d, const nh_config &c, uint32_t *max_delay) {
    auto fin = [](float x) { return std::isfinite(x); };
    if (!(c.dt_ms > 0.0f) || !fin(c.dt_ms)) return set_error(NH_ERR_ARG, "dt_ms must be positive");
    if (!(fin(c.synapse_tau_ms) && c.synapse_tau_ms >= 0.0f && fin(c.readout_tau_ms) &&
          c.readout_tau_ms >= 0.0f))
        return set_error(NH_ERR_ARG, "synapse and readout tau must be finite and >= 0");
    if (!(fin(c.noise_rate_hz) && c.noise_rate_hz >= 0.0f && fin(c.noise_kick_mv) && fin(c.noise_sd) &&
          c.noise_sd >= 0.0f))
        return set_error(NH_ERR_ARG, "noise rate, kick and sd must be finite (rate, sd >= 0)");
    if (c.int_weights && !(fin(c.int_scale) && c.int_scale > 0.0f))
        return set_error(NH_ERR_ARG, "int_scale must be positive and finite");
    if (d.n && (!d.src_ptr || !d.model || !d.p || !d.x0 || !d.model_ids))
        return set_error(NH_ERR_ARG, "network arrays missing");
    if (d.n_syn && (!d.tgt || !d.w || !d.delay)) return set_error(NH_ERR_ARG, "synapse arrays missing");
    if (d.n_models == 0 && d.n) return set_error(NH_ERR_ARG, "empty model table");
    if (d.n_models > 8) return set_error(NH_ERR_ARG, "model table too large");
    if (d.n && d.n_state != 2) return set_error(NH_ERR_ARG, "state row width must be 2");
    if (d.n_params > (uint32_t)kMaxParams) return set_error(NH_ERR_ARG, "param row too wide");
    if (d.n > 0x7FFFFFC0u) return set_error(NH_ERR_CAPACITY, "network exceeds 2^31 neurons");
    for (uint32_t m = 0; m < d.n_models; ++m)
        if (d.model_ids[m] > 2) return set_error(NH_ERR_ARG, "unknown model id " + std::to_string(d.model_ids[m]));
    // Per-neuron checks as reductions first (vectorised): the models in
    // use, whether any parameter or initial value has the all-ones exponent
    // of an infinity or NaN, whether src_ptr ever decreases; the loops that
    // name the offence run only when a reduction fails.
    uint32_t mmax = 0, used = 0;
    for (uint32_t i = 0; i < d.n; ++i) {
        mmax = std::max<uint32_t>(mmax, d.model[i]);
        used |= 1u << (d.model[i] & 31u);
    }
    bool bad_model = d.n && mmax >= d.n_models;
    for (uint32_t m = 0; !bad_model && m < d.n_models; ++m)
        if ((used >> m) & 1u)
            bad_model = (d.model_ids[m] == 1 && d.n_params < 10) || (d.model_ids[m] == 0 && d.n_params < 4) ||
                        (d.model_ids[m] == 2 && d.n_params < 5);
    for (uint32_t i = 0; bad_model && i < d.n; ++i) {
        if (d.model[i] >= d.n_models)
            return set_error(NH_ERR_ARG, "neuron " + std::to_string(i) + " names model " +
                                             std::to_string(d.model[i]) + " outside the model table");
        const uint8_t id = d.model_ids[d.model[i]];
        if (id == 1 && d.n_params < 10) return set_error(NH_ERR_ARG, "adex rows need 10 params");
        if (id == 0 && d.n_params < 4) return set_error(NH_ERR_ARG, "izh rows need 4 params");
        if (id == 2 && d.n_params < 5) return set_error(NH_ERR_ARG, "lif rows need 5 params");
    }
    auto all_finite = [](const float *v, uint64_t n) {
        uint32_t nonfin = 0;
        for (uint64_t q = 0; q < n; ++q) {
            uint32_t b;
            std::memcpy(&b, &v[q], 4);
            nonfin |= (b & 0x7F800000u) == 0x7F800000u;
        }
        return nonfin == 0;
    };
    if (!all_finite(d.p, (uint64_t)d.n * d.n_params)) return set_error(NH_ERR_ARG, "non-finite parameter");
    if (!all_finite(d.x0, (uint64_t)d.n * 2)) return set_error(NH_ERR_ARG, "non-finite initial state");
    if (d.n && d.src_ptr[0] != 0) return set_error(NH_ERR_ARG, "src_ptr[0] must be 0");
    uint32_t down = 0;
    for (uint32_t i = 0; i < d.n; ++i) down |= d.src_ptr[i + 1] < d.src_ptr[i];
    if (down) return set_error(NH_ERR_ARG, "src_ptr not monotone");
    const uint64_t E = d.n ? d.src_ptr[d.n] : 0;
    if (E != d.n_syn) return set_error(NH_ERR_ARG, "src_ptr[n] != n_syn");
    uint32_t maxd = 1;
    const double lim = 2147483647.0;
    // Fast pass: the synapse checks as reductions the compiler vectorises --
    // largest target, smallest and largest delay, largest |w| taken on its
    // bit pattern (monotone for non-negative floats; above the infinity
    // pattern only NaN) -- and the per-synapse pass below, which names the
    // first offending synapse, only when one of them fails.  The int-ring
    // bound is monotone in |w|, so its largest value decides it.
    bool need_slow = d.tau != nullptr;
    if (!need_slow && E) {
        uint32_t tmax = 0, wbits = 0;
        uint16_t dmin = 0xFFFF, dmax = 0;
        for (uint64_t j = 0; j < E; ++j) {
            tmax = std::max(tmax, d.tgt[j]);
            dmin = std::min(dmin, d.delay[j]);
            dmax = std::max(dmax, d.delay[j]);
            uint32_t b;
            std::memcpy(&b, &d.w[j], 4);
            wbits = std::max(wbits, b & 0x7FFFFFFFu);
        }
        float wmax;
        std::memcpy(&wmax, &wbits, 4);
        need_slow = tmax >= d.n || dmin < 1 || wbits >= 0x7F800000u ||
                    (c.int_weights && std::fabs((double)wmax * c.int_scale) >= lim);
        maxd = std::max<uint32_t>(maxd, dmax);
    }
    for (uint64_t j = 0; need_slow && j < E; ++j) {
        if (d.tgt[j] >= d.n) return set_error(NH_ERR_ARG, "synapse target out of range");
        if (d.delay[j] < 1) return set_error(NH_ERR_ARG, "delay must be >= 1 step");
        if (!fin(d.w[j])) return set_error(NH_ERR_ARG, "non-finite synapse weight");
        if (c.int_weights && std::fabs((double)d.w[j] * c.int_scale) >= lim)
            return set_error(NH_ERR_ARG, "weight x int_scale exceeds the int32 ring");
        if (d.tau && !(fin(d.tau[j]) && d.tau[j] >= 0.0f)) return set_error(NH_ERR_ARG, "synapse tau must be finite and >= 0");
        maxd = std::max<uint32_t>(maxd, d.delay[j]);
    }
    // Int ring: a slot of neuron i sums at most one arrival per in-synapse
    // (a synapse of delay d delivers into slot t only from a spike at t - d),
    // so the slot cannot wrap if the in-synapses' quantised magnitudes sum
    // below 2^31.  Refused otherwise: the CUDA int ring and the oracle would
    // both wrap silently.
    if (c.int_weights && E) {
        std::vector<int64_t> fan(d.n, 0);
        for (uint64_t j = 0; j < E; ++j)
            fan[d.tgt[j]] += std::llabs((long long)quantize(d.w[j], c.int_scale));
        for (uint32_t i = 0; i < d.n; ++i)
            if (fan[i] > 2147483647LL)
                return set_error(NH_ERR_NET, "neuron " + std::to_string(i) +
                                                 ": its in-synapses' weights x int_scale sum past the int32 "
                                                 "ring (a slot could overflow); lower int_scale");
    }
    if (d.n_inputs && d.stim_ptr) {
        if (!d.stim_input && d.stim_ptr[d.n]) return set_error(NH_ERR_ARG, "stimulus arrays missing");
        if (d.stim_ptr[0] != 0) return set_error(NH_ERR_ARG, "stim_ptr[0] must be 0");
        for (uint32_t i = 0; i < d.n; ++i) {
            if (d.stim_ptr[i + 1] < d.stim_ptr[i]) return set_error(NH_ERR_ARG, "stim_ptr not monotone");
            for (uint32_t j = d.stim_ptr[i]; j < d.stim_ptr[i + 1]; ++j) {
                if (d.stim_input[j] >= d.n_inputs) return set_error(NH_ERR_ARG, "stimulus input index out of range");
                if (!d.stim_w || !fin(d.stim_w[j])) return set_error(NH_ERR_ARG, "non-finite stimulus weight");
            }
        }
    }
    if (max_delay) *max_delay = maxd;
    return NH_OK;
}

int Arena::build(uint32_t nn, const nh_net_desc *nets, const nh_config *cfgs, uint32_t raster_steps,
                 const Tuning &tune) {
    if (nn == 0) return set_error(NH_ERR_ARG, "no networks");
    if (nn > 65535) return set_error(NH_ERR_ARG, "at most 65535 networks per arena");
    n_nets = nn;
    ring_int = cfgs[0].int_weights ? 1 : 0;
    slots.resize(nn);
    // ---- validate and lay out
    uint64_t off = 0;
    uint32_t maxd = 1;
    bool any_izh = false, any_adex = false, any_lif = false;
    uint32_t P = 4;
    uint64_t total_syn = 0;
    uint32_t cols = 0;
    for (uint32_t k = 0; k < nn; ++k) {
        const nh_net_desc &d = nets[k];
        if ((cfgs[k].int_weights ? 1 : 0) != ring_int)
            return set_error(NH_ERR_ARG, "all packed networks must share int_weights");
        uint32_t dk = 1;
        TRY(check_desc(d, cfgs[k], &dk));
        maxd = std::max(maxd, dk);
        for (uint32_t i = 0; i < d.n; ++i) {
            const uint8_t id = d.model_ids[d.model[i]];
            if (id == 0) any_izh = true;
            else if (id == 1) any_adex = true;
            else any_lif = true;
        }
        const uint64_t E = d.n_syn;
        P = std::max(P, d.n_params);
        total_syn += E;
        slots[k].base = (uint32_t)off;
        slots[k].n = d.n;
        slots[k].cfg = cfgs[k];
        slots[k].n_inputs = d.n_inputs;
        slots[k].stim_base = cols;
        cols += d.n_inputs;
        off += align64(d.n);
        if (off > 0x7FFFFFFFull) return set_error(NH_ERR_CAPACITY, "arena exceeds 2^31 neurons");
    }
    n_total = (uint32_t)off;
    if (n_total == 0) n_total = 0;
    words = n_total / 64;
    D = std::max<uint32_t>(maxd + 1, 2);  // ring_depth, state.rs:221-224
    // synapse indices are 32-bit on the device (and one value is reserved)
    if (total_syn >= 0xFFFF0000ull) return set_error(NH_ERR_CAPACITY, "arena exceeds 2^32 - 2^16 synapses");
    n_syn = total_syn;
    // 0: every neuron Izhikevich, 1: every neuron AdEx, 2: per-neuron
    // registry id (mixed models, or any LIF)
    mmode = any_lif ? 2 : (any_adex ? (any_izh ? 2 : 1) : 0);
    stim_cols = cols;
    packed = (n_total <= (1u << 27) && D <= 32) ? 1 : 0;

    HIPCHK(hipGetDevice(&device));
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // timing-only events: no system-scope fence when they are recorded (each
    // such fence writes back and invalidates the caches; three of them cost
    // a 20-step run of the 100k network 8.5 us of wall time)
    HIPCHK(hipEventCreateWithFlags(&ev0, hipEventDisableSystemFence));
    HIPCHK(hipEventCreateWithFlags(&ev1, hipEventDisableSystemFence));

    // ---- CSR over arena ids, rows sorted by (delay, target)
    std::vector<uint32_t> src_ptr(n_total + 1, 0);
    std::vector<uint32_t> syn(std::max<uint64_t>(total_syn, 1));
    std::vector<uint16_t> d16;
    if (!packed) d16.resize(std::max<uint64_t>(total_syn, 1));
    std::vector<float> wf;
    std::vector<int32_t> wq;
    if (ring_int) wq.resize(std::max<uint64_t>(total_syn, 1));
    else wf.resize(std::max<uint64_t>(total_syn, 1));
    {
        // Row starts first (every synapse without its own tau, in row
        // order), then the rows themselves, each sorted by (delay, target)
        // with ties in input order and written at its start.  Rows are
        // independent, so chunks of rows go to a few threads.
        uint64_t o = 0;
        for (uint32_t k = 0; k < nn; ++k) {
            const nh_net_desc &d = nets[k];
            const uint32_t b = slots[k].base;
            for (uint32_t i = 0; i < align64(d.n); ++i) {  // padding neurons: empty rows
                src_ptr[b + i] = (uint32_t)o;               // (the total is checked below)
                if (i >= d.n) continue;
                uint64_t c = d.src_ptr[i + 1] - d.src_ptr[i];
                if (d.tau)
                    for (uint32_t j = d.src_ptr[i]; j < d.src_ptr[i + 1]; ++j) c -= d.tau[j] > 0.0f ? 1 : 0;
                o += c;
            }
        }
        if (o > 0xFFFFFFFFull) return set_error(NH_ERR_CAPACITY, "more than 2^32 synapses");
        src_ptr[n_total] = (uint32_t)o;
        n_syn = o;
        struct Rows {
            uint32_t k, i0, i1;
        };
        std::vector<Rows> tasks;
        for (uint32_t k = 0; k < nn; ++k)
            for (uint32_t i = 0; i < nets[k].n; i += 2048)
                tasks.push_back({k, i, std::min<uint32_t>(nets[k].n, i + 2048)});
        std::atomic<size_t> next{0};
        auto work = [&] {
            // key: (delay << 32 | target, input index) -- a total order, the
            // stable (delay, target) order
            std::vector<std::pair<uint64_t, uint32_t>> key;
            for (size_t q; (q = next.fetch_add(1)) < tasks.size();) {
                const nh_net_desc &d = nets[tasks[q].k];
                const uint32_t b = slots[tasks[q].k].base;
                const float scale = cfgs[tasks[q].k].int_scale;
                for (uint32_t i = tasks[q].i0; i < tasks[q].i1; ++i) {
                    key.clear();
                    for (uint32_t j = d.src_ptr[i]; j < d.src_ptr[i + 1]; ++j)
                        if (!(d.tau && d.tau[j] > 0.0f))  // custom-tau: own rows below
                            key.emplace_back(((uint64_t)d.delay[j] << 32) | d.tgt[j], j);
                    if (key.size() > 32) {
                        std::sort(key.begin(), key.end());
                    } else {
                        for (size_t a = 1; a < key.size(); ++a) {  // short rows: insertion sort
                            const auto v = key[a];
                            size_t c = a;
                            for (; c && v < key[c - 1]; --c) key[c] = key[c - 1];
                            key[c] = v;
                        }
                    }
                    uint64_t w = src_ptr[b + i];
                    for (const auto &kv : key) {
                        const uint32_t j = kv.second, tg = b + d.tgt[j];
                        if (packed) syn[w] = tg | ((uint32_t)d.delay[j] << 27);
                        else {
                            syn[w] = tg;
                            d16[w] = d.delay[j];
                        }
                        if (ring_int) wq[w] = quantize(d.w[j], scale);
                        else wf[w] = d.w[j];
                        ++w;
                    }
                }
            }
        };
        run_workers(work, tasks.size());
    }
    // ---- custom-tau synapses (XSyn::build, sim/src/state.rs:245-274): slot k
    // in CSR order per network; per-source rows (k, delay, w) for the push,
    // target-major lists of k for the fold-in at the target
    std::vector<uint32_t> xs_lo(n_total, 0), xs_hi(n_total, 0), xs_k, xt_cnt(n_total + 1, 0), xtgt;
    std::vector<uint16_t> xs_d;
    std::vector<float> xs_wf, xdecay;
    std::vector<int32_t> xs_wq;
    for (uint32_t k = 0; k < nn; ++k) {
        const nh_net_desc &d = nets[k];
        if (!d.tau) continue;
        const uint32_t b = slots[k].base;
        const float dt = cfgs[k].dt_ms, scale = cfgs[k].int_scale;
        const uint32_t k0 = (uint32_t)xdecay.size();
        // slot numbering follows CSR order j
        std::vector<uint32_t> slot_of(d.n_syn, 0xFFFFFFFFu);
        for (uint64_t j = 0; j < d.n_syn; ++j)
            if (d.tau[j] > 0.0f) {
                slot_of[j] = (uint32_t)xdecay.size();
                const float tau = d.tau[j] > 1e-6f ? d.tau[j] : 1e-6f;
                xdecay.push_back(std::exp(-dt / tau));
                xtgt.push_back(b + d.tgt[j]);
                xt_cnt[b + d.tgt[j] + 1]++;
            }
        (void)k0;
        for (uint32_t i = 0; i < d.n; ++i) {
            xs_lo[b + i] = (uint32_t)xs_k.size();
            for (uint32_t j = d.src_ptr[i]; j < d.src_ptr[i + 1]; ++j) {
                if (slot_of[j] == 0xFFFFFFFFu) continue;
                xs_k.push_back(slot_of[j]);
                xs_d.push_back(d.delay[j]);
                if (ring_int) xs_wq.push_back(quantize(d.w[j], scale));
                else xs_wf.push_back(d.w[j]);
            }
            xs_hi[b + i] = (uint32_t)xs_k.size();
        }
        for (uint32_t i = d.n; i < align64(d.n); ++i) xs_lo[b + i] = xs_hi[b + i] = (uint32_t)xs_k.size();
    }
    n_x = (uint32_t)xdecay.size();
    std::vector<uint32_t> xt_ptr, xt_idx;
    if (n_x) {
        xt_ptr.assign(n_total + 1, 0);
        for (uint32_t i = 0; i < n_total; ++i) xt_ptr[i + 1] = xt_ptr[i] + xt_cnt[i + 1];
        xt_idx.resize(n_x);
        std::vector<uint32_t> fill(n_total, 0);
        for (uint32_t kx = 0; kx < n_x; ++kx) xt_idx[xt_ptr[xtgt[kx]] + fill[xtgt[kx]]++] = kx;
    }
    // total out-degree per neuron (events on readback)
    h_deg_.assign(n_total, 0);
    for (uint32_t i = 0; i < n_total; ++i)
        h_deg_[i] = (src_ptr[i + 1] - src_ptr[i]) + (n_x ? xs_hi[i] - xs_lo[i] : 0);
    // one kernel per step (own-spike push) unless the two-kernel or the
    // paired shape is asked for
    fused = tune.fused == 0 ? 0 : (tune.fused == 2 || tune.fused == 3 ? tune.fused : 1);
    // Delay-1 in-tables for the paired shape (asked for, or left to the
    // autotuner on an arena small enough to be co-resident).  The ring then
    // gets one slot more than the deepest delay needs, so the pushes of a
    // paired launch never reach the two slots it reads.
    pair_tables = (tune.fused == 2 || tune.fused == 3 || (tune.fused == -1 && n_total <= (1u << 20))) && n_x == 0 &&
                  n_total > 0 && n_total < (1u << 26);
    std::vector<uint32_t> pl_ptr, pl_ent;
    std::vector<int32_t> pl_wq;
    std::vector<float> pl_wf;
    if (pair_tables) {
        D += 1;
        // rows are sorted by delay, so a row's delay-1 synapses are its prefix
        auto tgt_of = [&](uint64_t o) { return packed ? (syn[o] & 0x07FFFFFFu) : syn[o]; };
        auto d1 = [&](uint64_t o) { return (packed ? syn[o] >> 27 : (uint32_t)d16[o]) == 1u; };
        pl_ptr.assign((size_t)words + 1, 0);
        for (uint32_t i = 0; i < n_total; ++i)
            for (uint64_t o = src_ptr[i]; o < src_ptr[i + 1] && d1(o); ++o) pl_ptr[tgt_of(o) / 64 + 1]++;
        for (uint32_t w = 0; w < words; ++w) pl_ptr[w + 1] += pl_ptr[w];
        const uint32_t E1 = pl_ptr[words];
        pl_ent.assign(std::max<uint32_t>(E1, 1), 0);
        if (ring_int) pl_wq.assign(std::max<uint32_t>(E1, 1), 0);
        else pl_wf.assign(std::max<uint32_t>(E1, 1), 0.0f);
        // auto: paired where the tables are small (at most four pull rounds
        // of 64 entries per wave); the launcher still checks co-residency
        uint32_t most = 0;
        for (uint32_t w = 0; w < words; ++w) most = std::max(most, pl_ptr[w + 1] - pl_ptr[w]);
        if (tune.fused == -1 && most <= 4u * 64u) fused = 2;
        std::vector<uint32_t> fill(pl_ptr.begin(), pl_ptr.end() - 1);
        // entries of a wave in source order
        for (uint32_t i = 0; i < n_total; ++i)
            for (uint64_t o = src_ptr[i]; o < src_ptr[i + 1] && d1(o); ++o) {
                const uint32_t tg = tgt_of(o), e = fill[tg / 64]++;
                pl_ent[e] = i | ((tg & 63u) << 26);
                if (ring_int) pl_wq[e] = wq[o];
                else pl_wf[e] = wf[o];
            }
    }

    // ---- parameters: type table when the arena has <= 256 distinct rows
    std::map<std::vector<uint32_t>, uint32_t> rows;
    std::vector<uint8_t> prow(n_total, 0);
    std::vector<PRow> ptab;
    bool table_ok = true;
    for (uint32_t k = 0; k < nn && table_ok; ++k) {
        const nh_net_desc &d = nets[k];
        for (uint32_t i = 0; i < d.n; ++i) {
            std::vector<uint32_t> key(kMaxParams + 1, 0);
            key[0] = d.model_ids[d.model[i]];
            for (uint32_t q = 0; q < d.n_params; ++q)
                std::memcpy(&key[1 + q], &d.p[(uint64_t)i * d.n_params + q], 4);
            auto it = rows.find(key);
            uint32_t r;
            if (it == rows.end()) {
                if (rows.size() >= 256) {
                    table_ok = false;
                    break;
                }
                r = (uint32_t)rows.size();
                rows.emplace(key, r);
                PRow pr{};
                for (int q = 0; q < kMaxParams; ++q) std::memcpy(&pr.p[q], &key[1 + q], 4);
                pr.model = key[0];
                ptab.push_back(pr);
            } else {
                r = it->second;
            }
            prow[slots[k].base + i] = (uint8_t)r;
        }
    }
    // <= 4 rows ride in the kernel arguments, <= 256 in a global table,
    // otherwise per-neuron rows
    pmode = !table_ok ? 1 : (ptab.size() <= (size_t)kKargRows ? 0 : 2);
    krows_ = ptab;
    std::vector<float> pflat;
    std::vector<uint8_t> mreg;
    if (pmode == 1) {
        pflat.assign((uint64_t)n_total * P, 0.0f);
        mreg.assign(n_total, 0);
        for (uint32_t k = 0; k < nn; ++k) {
            const nh_net_desc &d = nets[k];
            for (uint32_t i = 0; i < d.n; ++i) {
                const uint32_t a = slots[k].base + i;
                mreg[a] = d.model_ids[d.model[i]];
                for (uint32_t q = 0; q < d.n_params; ++q)
                    pflat[(uint64_t)a * P + q] = d.p[(uint64_t)i * d.n_params + q];
            }
        }
        prow.clear();
    }
    this->P_ = P;

    // ---- per-network records, per-wave network index
    recs.resize(nn);
    std::vector<uint32_t> wave_net(std::max<uint32_t>(words, 1), 0);
    for (uint32_t k = 0; k < nn; ++k) {
        const nh_config &c = cfgs[k];
        NetRec &r = recs[k];
        std::memset(&r, 0, sizeof(r));
        r.base = slots[k].base;
        r.n = slots[k].n;
        r.seed = c.seed;
        r.t0 = 0;
        fill_noise(r, c);
        r.int_scale = c.int_scale;
        r.stim_base = slots[k].stim_base;
        r.active = 1;
        r.dt = c.dt_ms;
        for (uint32_t w = r.base / 64; w < (r.base + align64(r.n)) / 64; ++w) wave_net[w] = k;
    }

    // ---- initial state and stimulus map over arena ids
    std::vector<float2> x0(std::max<uint32_t>(n_total, 1), make_float2(0.f, 0.f));
    std::vector<uint32_t> stim_ptr;
    std::vector<uint32_t> stim_input;
    std::vector<float> stim_w;
    bool any_stim = false;
    for (uint32_t k = 0; k < nn; ++k)
        if (nets[k].n_inputs && nets[k].stim_ptr && nets[k].stim_ptr[nets[k].n]) any_stim = true;
    if (any_stim) stim_ptr.assign(n_total + 1, 0);
    for (uint32_t k = 0; k < nn; ++k) {
        const nh_net_desc &d = nets[k];
        const uint32_t b = slots[k].base;
        for (uint32_t i = 0; i < d.n; ++i)
            x0[b + i] = make_float2(d.x0[(uint64_t)i * 2], d.x0[(uint64_t)i * 2 + 1]);
        if (!any_stim) continue;
        for (uint32_t i = 0; i < align64(d.n); ++i) {
            stim_ptr[b + i] = (uint32_t)stim_input.size();
            if (i < d.n && d.n_inputs && d.stim_ptr)
                for (uint32_t j = d.stim_ptr[i]; j < d.stim_ptr[i + 1]; ++j) {
                    stim_input.push_back(d.stim_input[j]);
                    stim_w.push_back(d.stim_w[j]);
                }
        }
    }
    if (any_stim) stim_ptr[n_total] = (uint32_t)stim_input.size();

    // ---- device buffers
    TRY(alloc(d_t, 8));
    TRY(upload(d_nets, recs));
    TRY(upload(d_wave_net, wave_net));
    TRY(upload(d_x0, x0));
    TRY(upload(d_x, x0));
    TRY(alloc(d_s, (size_t)n_total * 4));
    TRY(alloc(d_ring, (size_t)D * n_total * 4));
    TRY(alloc(d_ext, (size_t)n_total * 4));
    if (pmode == 0 || pmode == 2) {
        // row index per neuron + the rows (mode 0 also carries the rows in
        // the kernel arguments; mode 2 reads them from this table)
        TRY(upload(d_prow, prow));
        TRY(upload(d_ptab, ptab));
    } else {
        TRY(upload(d_p, pflat));
        TRY(upload(d_mreg, mreg));
    }
    if (any_stim) {
        TRY(upload(d_stim_ptr, stim_ptr));
        TRY(upload(d_stim_input, stim_input));
        TRY(upload(d_stim_w, stim_w));
    }
    stim_group.resize(nn);
    for (uint32_t k = 0; k < nn; ++k) stim_group[k] = make_uint2(slots[k].stim_base, slots[k].n_inputs);
    TRY(alloc_stim(1024));
    TRY(alloc(d_spk, (size_t)3 * std::max<uint32_t>(n_total, 1) * 4));
    TRY(alloc(d_cnt, 4 * 4));
    bits_rows = 0;
    if (cfgs[0].record_spikes) {
        bits_rows = std::max<uint32_t>(cfgs[0].spike_record_steps, 2);
        TRY(alloc(d_bits, (size_t)bits_rows * std::max<uint32_t>(words, 1) * 8));
    }
    TRY(alloc(d_pn, (size_t)n_total * 4));
    TRY(alloc(d_pn_mark, (size_t)n_total * 4));
    TRY(alloc(d_last, (size_t)n_total * 4));
    HIPCHK(hipMemset(d_last.p, 0xFF, d_last.bytes));
    TRY(alloc(d_net_spikes, (size_t)nn * 8));
    TRY(upload(d_src_ptr, src_ptr));
    TRY(upload(d_syn, syn));
    if (!packed) TRY(upload(d_delay16, d16));
    if (ring_int) TRY(upload(d_w, wq));
    else TRY(upload(d_w, wf));
    if (pair_tables) {
        TRY(upload(d_pl_ptr, pl_ptr));
        TRY(upload(d_pl_ent, pl_ent));
        if (ring_int) TRY(upload(d_pl_w, pl_wq));
        else TRY(upload(d_pl_w, pl_wf));
        TRY(alloc(d_pub, (size_t)words * 16));
        TRY(alloc(d_pair_abort, 4));
        HIPCHK(h_pair_abort.alloc());
        TRY(alloc(d_pub3, (size_t)words * 16 * kPubDepth));
        TRY(alloc(d_gctr, 4 * 64 * 4));
        TRY(alloc(d_done, 4 * 64 * 4));
    }
    TRY(alloc(d_totals, 8 * 8));
    if (n_x) {
        TRY(upload(d_xt_ptr, xt_ptr));
        TRY(upload(d_xt_idx, xt_idx));
        TRY(upload(d_xdecay, xdecay));
        TRY(alloc(d_xsyn, (size_t)n_x * 4));
        TRY(alloc(d_xring, (size_t)D * n_x * 4));
        TRY(upload(d_xs_lo, xs_lo));
        TRY(upload(d_xs_hi, xs_hi));
        TRY(upload(d_xs_k, xs_k));
        TRY(upload(d_xs_d, xs_d));
        if (ring_int) TRY(upload(d_xs_w, xs_wq));
        else TRY(upload(d_xs_w, xs_wf));
        xargs_h.n_x = n_x;
        xargs_h.xt_ptr = d_xt_ptr.as<uint32_t>();
        xargs_h.xt_idx = d_xt_idx.as<uint32_t>();
        xargs_h.xdecay = d_xdecay.as<float>();
        xargs_h.xsyn = d_xsyn.as<float>();
        xargs_h.xring = d_xring.p;
        xargs_h.xs_lo = d_xs_lo.as<uint32_t>();
        xargs_h.xs_hi = d_xs_hi.as<uint32_t>();
        xargs_h.xs_k = d_xs_k.as<uint32_t>();
        xargs_h.xs_d = d_xs_d.as<uint16_t>();
        xargs_h.xs_w = d_xs_w.p;
        TRY(alloc(d_xargs, sizeof(XArgs)));
        HIPCHK(hipMemcpy(d_xargs.p, &xargs_h, sizeof(XArgs), hipMemcpyHostToDevice));
    }
    // raster window: one spike word per wave per step (no atomics; the
    // host decodes ids, ascending, when it drains the window)
    rast_steps = raster_steps;
    if (rast_steps) TRY(alloc(d_rast_bits, (size_t)rast_steps * std::max<uint32_t>(words, 1) * 8));
    TRY(alloc(d_rast_ctl, 4 * 8));
    h_src_ptr_ = std::move(src_ptr);
    HIPCHK(hipDeviceSynchronize());

    if (tune.block) {
        if (tune.block % 64 || tune.block > (uint32_t)kMaxBlock)
            return set_error(NH_ERR_ARG, "block must be a multiple of 64 up to 512");
        block = tune.block;
    }
    if (tune.lds_bins) {
        const uint32_t b = tune.lds_bins;
        if (b < 256u || b > 8192u || (b & (b - 1u)))
            return set_error(NH_ERR_ARG, "lds_bins must be 0 or a power of two in 256..8192");
        if ((uint64_t)D * n_total >= 0xFFFFFFFFull)
            return set_error(NH_ERR_ARG, "lds_bins: the ring is too large for 32-bit bin keys");
        lds_bins = b;
    }
    use_graph = tune.use_graph >= 0 ? tune.use_graph : 1;
    graph_steps_req = tune.graph_steps ? tune.graph_steps : 64;
    graph_steps = 0;
    scatter_blocks = tune.scatter_blocks;
    return NH_OK;
}

void Arena::fill_noise(NetRec &r, const nh_config &c) {
    // sim/src/cpu.rs:197-204 (host-side scalars, f32)
    r.pkick = c.noise_rate_hz * c.dt_ms / 1000.0f;
    r.kick = c.noise_kick_mv != 0.0f ? c.noise_kick_mv / c.dt_ms : 0.0f;
    r.sd = c.noise_sd;
    const float tau = c.synapse_tau_ms > 1e-6f ? c.synapse_tau_ms : 1e-6f;
    r.decay_s = std::exp(-c.dt_ms / tau);
}

int Arena::upload_recs() {
    HIPCHK(hipMemcpyAsync(d_nets.p, recs.data(), recs.size() * sizeof(NetRec),
                          hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return NH_OK;
}

StepArgs Arena::args() const {
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n_total = n_total;
    a.block = block;
    a.D = D;
    a.t_dev = d_t.as<uint64_t>();
    a.self = d_args.as<StepArgs>();
    a.k = 0;
    a.slot = 0;
    a.single_net = (n_nets == 1 && !dynamic) ? 1 : 0;
    a.wave_net = d_wave_net.as<uint32_t>();
    a.nets = d_nets.as<NetRec>();
    a.net0 = recs[0];
    a.x = d_x.as<float2>();
    a.s = d_s.as<float>();
    a.ring = d_ring.p;
    {
        const uint64_t rb = (uint64_t)D * n_total * 4;
        a.ring_buf = rb < 0x80000000ull ? 1u : 0u;   // the dropped-slot offset must be out of range
        a.ring_bytes = a.ring_buf ? (uint32_t)rb : 0u;
    }
    a.ext = has_ext ? d_ext.as<float>() : nullptr;
    a.n_krows = pmode == 0 ? (uint32_t)krows_.size() : 0;
    for (size_t r = 0; r < krows_.size() && r < (size_t)kKargRows; ++r) a.krow[r] = krows_[r];
    a.prow = d_prow.as<uint8_t>();
    a.ptab = d_ptab.as<PRow>();
    a.p = d_p.as<float>();
    a.mreg = d_mreg.as<uint8_t>();
    a.P = P_;
    a.stim_lo = d_stim_ptr.as<uint32_t>();
    a.stim_hi = d_stim_hi.p ? d_stim_hi.as<uint32_t>()
                            : (d_stim_ptr.p ? d_stim_ptr.as<uint32_t>() + 1 : nullptr);
    a.stim_input = d_stim_input.as<uint32_t>();
    a.stim_w = d_stim_w.as<float>();
    a.stim_win = d_stim_win.as<float>();
    a.stim_nz = d_stim_nz.as<uint32_t>();
    a.stim_groups = (uint32_t)stim_group.size();
    a.stim_cols = stim_cols;
    a.spk = d_spk.as<uint32_t>();
    a.cnt = d_cnt.as<uint32_t>();
    a.bits = bits_rows ? d_bits.as<uint64_t>() : nullptr;
    a.bits_rows = bits_rows ? bits_rows : 1;
    a.words = words;
    a.pn_total = d_pn.as<uint32_t>();
    a.last_spike = d_last.as<uint32_t>();
    a.net_spikes = (n_nets > 1 || dynamic) ? d_net_spikes.as<uint64_t>() : nullptr;
    a.rast_bits = rast_steps ? d_rast_bits.as<uint64_t>() : nullptr;
    a.rast_ctl = d_rast_ctl.as<uint64_t>();
    a.rast_cap_steps = rast_steps;
    if (d_row_hi.p) {
        a.row_lo = d_src_ptr.as<uint32_t>();
        a.row_hi = d_row_hi.as<uint32_t>();
    } else {
        a.row_lo = d_src_ptr.as<uint32_t>();
        a.row_hi = d_src_ptr.as<uint32_t>() + 1;
    }
    a.wcnt = d_wcnt.as<uint32_t>();
    a.n_win = n_win;
    a.xt = n_x ? d_xargs.as<XArgs>() : nullptr;
    a.syn = d_syn.as<uint32_t>();
    a.delay16 = d_delay16.as<uint16_t>();
    a.w = d_w.p;
    a.totals = d_totals.as<uint64_t>();
    a.watch = d_watch.as<uint32_t>();
    a.n_watch = n_watch;
    a.pl_ptr = d_pl_ptr.as<uint32_t>();
    a.pl_hi = d_pl_hi.p ? d_pl_hi.as<uint32_t>() : (d_pl_ptr.p ? d_pl_ptr.as<uint32_t>() + 1 : nullptr);
    a.pl_ent = d_pl_ent.as<uint32_t>();
    a.pl_w = d_pl_w.p;
    a.pub = d_pub.as<uint64_t>();
    a.pair_abort = d_pair_abort.as<uint32_t>();
    a.pair_abort_host = h_pair_abort.dev;
    a.pair_spins = pair_spins ? pair_spins : kPairSpinsDefault;
    a.pair_mute = pair_mute;
    a.pub3 = d_pub3.as<uint64_t>();
    a.gctr = d_gctr.as<uint32_t>();
    a.done = d_done.as<uint32_t>();
    a.n_groups = run_groups;
    a.run_steps = 1;
    a.bin_entries = binned() ? lds_bins : 0u;
    a.xrec = d_xrec.as<float2>();
    a.xrec_depth = xrec_depth;
    return a;
}

uint32_t Arena::neuron_blocks() const {
    const uint32_t n = grid_limit ? std::min(grid_limit, n_total) : n_total;
    return std::max<uint32_t>((n + block - 1) / block, 1);
}

uint32_t Arena::scat_blocks() const {
    if (scatter_blocks) return scatter_blocks;
    return std::min<uint32_t>(std::max<uint32_t>(neuron_blocks(), 1), 2048 * 256 / block);
}

// Every pointer a kernel variant dereferences must be set for that
// variant; a mismatch between the build and the launch mode is reported
// here instead of reaching the device.
int Arena::check_args(const StepArgs &a) const {
    auto need = [](const void *p, const char *what) -> int {
        return p ? NH_OK : set_error(NH_ERR_ARG, std::string("step argument missing: ") + what);
    };
    TRY(need(a.t_dev, "t_dev"));
    TRY(need(a.self, "self"));
    TRY(need(a.x, "x"));
    TRY(need(a.s, "s"));
    TRY(need(a.ring, "ring"));
    TRY(need(a.pn_total, "pn_total"));
    TRY(need(a.last_spike, "last_spike"));
    TRY(need(a.row_lo, "row_lo"));
    TRY(need(a.row_hi, "row_hi"));
    TRY(need(a.spk, "spk"));
    TRY(need(a.cnt, "cnt"));
    TRY(need(a.totals, "totals"));
    TRY(need(a.rast_ctl, "rast_ctl"));
    if (!a.single_net) {
        TRY(need(a.wave_net, "wave_net"));
        TRY(need(a.nets, "nets"));
    }
    if (n_syn) {
        TRY(need(a.syn, "syn"));
        TRY(need(a.w, "w"));
        if (!packed) TRY(need(a.delay16, "delay16"));
    }
    if (pmode == 0 && a.n_krows > 1) TRY(need(a.prow, "prow"));
    if (pmode == 0 && (a.n_krows < 1 || a.n_krows > (uint32_t)kKargRows))
        return set_error(NH_ERR_ARG, "kernel-argument parameter rows out of range");
    if (pmode == 2) {
        TRY(need(a.prow, "prow"));
        TRY(need(a.ptab, "ptab"));
    }
    if (pmode == 1) {
        TRY(need(a.p, "p"));
        if (mmode == 2) TRY(need(a.mreg, "mreg"));
    }
    if (a.stim_lo) {
        TRY(need(a.stim_hi, "stim_hi"));
        TRY(need(a.stim_input, "stim_input"));
        TRY(need(a.stim_w, "stim_w"));
        TRY(need(a.stim_win, "stim_win"));
        TRY(need(a.stim_nz, "stim_nz"));
    }
    if (a.xt) {
        const XArgs &x = xargs_h;
        if (x.n_x != n_x) return set_error(NH_ERR_ARG, "custom-tau arguments out of date");
        TRY(need(x.xt_ptr, "xt_ptr"));
        TRY(need(x.xt_idx, "xt_idx"));
        TRY(need(x.xdecay, "xdecay"));
        TRY(need(x.xsyn, "xsyn"));
        TRY(need(x.xring, "xring"));
        TRY(need(x.xs_lo, "xs_lo"));
        TRY(need(x.xs_hi, "xs_hi"));
        TRY(need(x.xs_k, "xs_k"));
        TRY(need(x.xs_d, "xs_d"));
        TRY(need(x.xs_w, "xs_w"));
    }
    if (a.n_watch) {
        TRY(need(a.watch, "watch"));
        TRY(need(a.xrec, "xrec"));
    }
    if (fused == 3 && pair_tables) {
        TRY(need(a.pub3, "pub3"));
        TRY(need(a.gctr, "gctr"));
        TRY(need(a.done, "done"));
        if (a.n_groups == 0 || a.n_groups > 64 || a.n_groups > neuron_blocks())
            return set_error(NH_ERR_ARG, "bad workgroup group count");
    }
    if (a.pl_ptr) {
        TRY(need(a.pl_hi, "pl_hi"));
        TRY(need(a.pl_ent, "pl_ent"));
        TRY(need(a.pl_w, "pl_w"));
        TRY(need(a.pub, "pub"));
        TRY(need(a.pair_abort, "pair_abort"));
    }
    if (a.block % 64 || a.block == 0 || a.block > (uint32_t)kMaxBlock)
        return set_error(NH_ERR_ARG, "bad workgroup size");
    return NH_OK;
}

// The persistent kernel waits on every workgroup of its launch, so it runs
// only when the whole grid fits on the device at once (at most half of the
// waves the device holds for it, as for the paired kernel) and for the lean
// case it covers: one network, no stimulus, no custom-tau synapses, no
// window counters, no watched neurons.
bool Arena::run_usable() {
    if (fused != 3 || !pair_tables || !lean() || n_watch || n_x || words == 0 || !d_pub3.p) return false;
    if (run_cap_block != block) {
        run_cap_waves = run_resident_waves(ring_int, pmode, mmode, packed != 0, block);
        run_cap_block = block;
    }
    const uint64_t need = (uint64_t)neuron_blocks() * (block / 64) * 2;
    if (need > run_cap_waves) return false;
    // groups of workgroups for the done tags: at most 64 (one poll lane each)
    run_groups = std::min<uint32_t>(32u, neuron_blocks());
    return true;
}

// The paired kernel waits on other waves of its own launch, so it runs only
// when the whole grid fits on the device at once -- here: when it needs at
// most half of the waves the device holds for that kernel and workgroup size
// -- and only for the features it covers (no watched neurons, no custom-tau
// synapses).  A single network whose grid is too large for k_pair may still
// fit k_pair8, the build held to 8 waves per SIMD (pair_w8).
bool Arena::pair_usable() {
    pair_w8 = false;
    if (fused != 2 || !pair_tables || n_watch || n_x || words == 0) return false;
    if (D > 65536u) return false;   // the paired kernel takes its two ring slots as 16-bit halves
    if (pair_cap_block != block) {
        pair_cap_waves = pair_resident_waves(ring_int, pmode, mmode, packed != 0, lean(), false, block);
        pair_cap8_waves = lean() ? pair_resident_waves(ring_int, pmode, mmode, packed != 0, true, true, block) : 0;
        pair_cap_block = block;
    }
    const uint64_t need = (uint64_t)neuron_blocks() * (block / 64) * 2;
    if (need <= pair_cap_waves) return true;
    pair_w8 = lean() && need <= pair_cap8_waves;
    return pair_w8;
}

// A wave whose bounded wait ran out raises the device flag.  In a paired
// launch the wave then recomputed the sources it waited for (results
// unchanged): the give-up is counted, not reported.  A persistent launch
// cannot recompute (its sources' state lives in registers), so there the
// state is invalid: reported (once) at the next synchronisation.
int Arena::check_pair_abort() {
    if (!pair_launched || !d_pair_abort.p) return NH_OK;
    pair_launched = false;
    const bool persist = persist_launched;
    persist_launched = false;
    // the host mirror (mapped memory, written by the wave that gave up) is
    // read directly: no copy on the common path
    if (!h_pair_abort.get()) return NH_OK;
    h_pair_abort.clear();
    HIPCHK(hipMemsetAsync(d_pair_abort.p, 0, 4, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (!persist) {
        ++pair_fallbacks;
        return NH_OK;
    }
    return set_error(NH_ERR_HIP, "persistent launch: a wave's wait for its sources ran out "
                                 "(workgroups not co-resident); the state is invalid, reset the simulation");
}

// Bring the step counter d_t to `to` (an advance kernel enqueued; nothing
// if it is there already).
int Arena::sync_t_dev(uint64_t to) {
    if (to < t_dev_val) return set_error(NH_ERR_ARG, "step counter ahead of the simulation");
    while (t_dev_val != to) {
        const uint32_t d = (uint32_t)std::min<uint64_t>(to - t_dev_val, 1ull << 31);
        launch_advance(d_t.as<uint64_t>(), d, stream);
        t_dev_val += d;
    }
    HIPCHK(hipGetLastError());
    return NH_OK;
}

// Enqueue k steps starting at arena step t0.  The kernels read the absolute
// step as d_t + a.k.  A direct chunk without stimulus inputs gives a.k as
// the offset of each step from the counter's value (rebased when the offset
// would pass kMaxChunkSteps, the most the one-step kernel's packed argument
// holds), so it needs no counter update of its own: an advance kernel per
// run cost 11 us of wall time in a 20-step run of the 100k network (75 -> 63
// us).  Otherwise (stimulus rows are indexed by a.k within the chunk, and a
// captured chunk replays from any step) d_t is brought to t0 first and
// advanced by k after the chunk.
int Arena::launch_chunk(uint64_t t0, uint32_t k) {
    if (k > kMaxChunkSteps) return set_error(NH_ERR_ARG, "launch chunk longer than the step kernels take");
    StepArgs a = args();
    TRY(check_args(a));
    const bool offset = !capturing && !d_stim_ptr.p;
    if (offset) {
        if (t0 < t_dev_val || t0 + k - t_dev_val > kMaxChunkSteps) TRY(sync_t_dev(t0));
    } else if (!capturing) {
        TRY(sync_t_dev(t0));
    }
    const uint32_t base = offset ? (uint32_t)(t0 - t_dev_val) : 0u;
    const uint32_t nb = neuron_blocks(), sb = scat_blocks();
    const bool pair = pair_usable();
    for (uint32_t j = 0; j < k;) {
        a.k = base + j;
        a.slot = (uint32_t)((t0 + j) % D);
        if (pair && j + 1 < k) {
            launch_pair(a, nb, ring_int, pmode, mmode, packed != 0, lean(), pair_w8, stream);
            TRY(note_launches(1));
            j += 2;
            continue;
        }
        if (binned()) launch_step_binned(a, nb, ring_int, pmode, mmode, packed != 0, stream);
        else launch_step(a, nb, ring_int, pmode, mmode, fused != 0, packed != 0, lean_step(), stream);
        if (!fused) launch_scatter(a, sb, ring_int, packed != 0, stream);
        if (n_watch) launch_watch(a, stream);
        TRY(note_launches(1));
        ++j;
    }
    if (!offset) {
        launch_advance(d_t.as<uint64_t>(), k, stream);
        if (!capturing) t_dev_val += k;
    }
    HIPCHK(hipGetLastError());
    return NH_OK;
}

// After a run's step launches (not while a graph is captured): the first
// one records ev_tail behind itself, later ones are counted.
int Arena::note_launches(uint64_t n) {
    if (!tail_marking) return NH_OK;
    if (!tail_marked) {
        HIPCHK(hipEventRecord(ev_tail, stream));
        tail_marked = true;
    } else {
        tail_launches += n;
    }
    return NH_OK;
}

// One persistent launch of k steps from arena step t0 (run_usable() true),
// then the chunk advance of the device step counter.
int Arena::launch_persistent(uint64_t t0, uint32_t k) {
    StepArgs a = args();
    a.k = 0;
    a.slot = (uint32_t)(t0 % D);
    a.run_steps = k;
    TRY(check_args(a));
    TRY(sync_t_dev(t0));
    launch_run(a, neuron_blocks(), ring_int, pmode, mmode, packed != 0, stream);
    TRY(note_launches(1));
    launch_advance(d_t.as<uint64_t>(), k, stream);
    t_dev_val += k;
    HIPCHK(hipGetLastError());
    return NH_OK;
}

int Arena::profile(uint64_t n_steps, double &main_ms, uint64_t &main_n, double &aux_ms,
                   uint64_t &aux_n) {
    main_ms = aux_ms = 0.0;
    main_n = aux_n = 0;
    if (n_total == 0 || n_steps == 0) {
        t += n_steps;
        return NH_OK;
    }
    TRY(sync_dev_args());
    constexpr uint32_t kPool = 64;  // steps per event batch
    std::vector<hipEvent_t> ev(4 * kPool, nullptr);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    int rc = NH_OK;
    const uint32_t nb = neuron_blocks(), sb = scat_blocks();
    while (n_steps && rc == NH_OK) {
        uint64_t k = std::min<uint64_t>(n_steps, std::min<uint32_t>(kPool, stim_cap));
        if (rast_steps) {
            const uint64_t room = rast_steps - (t - rast_step0);
            if (room == 0) {
                rc = drain_raster();
                continue;
            }
            k = std::min<uint64_t>(k, room);
        }
        if ((rc = upload_stim(t, (uint32_t)k))) break;
        if (run_usable()) {
            // one persistent launch of the k steps: one main launch
            (void)hipEventRecord(ev[0], stream);
            if ((rc = launch_persistent(t, (uint32_t)k))) break;
            (void)hipEventRecord(ev[1], stream);
            pair_launched = true;
            persist_launched = true;
            hipError_t e = hipStreamSynchronize(stream);
            if (e != hipSuccess) {
                rc = set_error(NH_ERR_HIP, std::string("profile: ") + hipGetErrorString(e));
                break;
            }
            if ((rc = check_pair_abort())) break;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) {
                main_ms += ms;
                ++main_n;
            }
            t += k;
            n_steps -= k;
            continue;
        }
        StepArgs a = args();
        if ((rc = check_args(a))) break;
        if ((rc = sync_t_dev(t))) break;
        // paired launches count as one main launch of two steps
        const bool pair = pair_usable();
        std::vector<uint32_t> firsts;
        for (uint32_t j = 0; j < k;) {
            a.k = j;
            a.slot = (uint32_t)((t + j) % D);
            firsts.push_back(j);
            (void)hipEventRecord(ev[4 * j], stream);
            if (pair && j + 1 < k) {
                launch_pair(a, nb, ring_int, pmode, mmode, packed != 0, lean(), pair_w8, stream);
                (void)hipEventRecord(ev[4 * j + 1], stream);
                pair_launched = true;
                j += 2;
                continue;
            }
            if (binned()) launch_step_binned(a, nb, ring_int, pmode, mmode, packed != 0, stream);
            else launch_step(a, nb, ring_int, pmode, mmode, fused != 0, packed != 0, lean_step(), stream);
            (void)hipEventRecord(ev[4 * j + 1], stream);
            if (!fused) {
                (void)hipEventRecord(ev[4 * j + 2], stream);
                launch_scatter(a, sb, ring_int, packed != 0, stream);
                (void)hipEventRecord(ev[4 * j + 3], stream);
            }
            if (n_watch) launch_watch(a, stream);
            ++j;
        }
        launch_advance(d_t.as<uint64_t>(), (uint32_t)k, stream);
        t_dev_val += k;
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            rc = set_error(NH_ERR_HIP, std::string("profile: ") + hipGetErrorString(e));
            break;
        }
        if ((rc = check_pair_abort())) break;
        for (uint32_t j : firsts) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[4 * j], ev[4 * j + 1]) == hipSuccess) {
                main_ms += ms;
                ++main_n;
            }
            if (!fused && hipEventElapsedTime(&ms, ev[4 * j + 2], ev[4 * j + 3]) == hipSuccess) {
                aux_ms += ms;
                ++aux_n;
            }
        }
        t += k;
        n_steps -= k;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}

// A captured chunk covers graph_steps steps, a multiple of the ring depth,
// and is only replayed from arena steps t = 0 (mod D): the ring slot of
// every step is then fixed per graph node.
int Arena::capture_graph() {
    tail_marking = false;   // nothing is recorded into the captured chunk
    if (graph_exec) {
        (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
    }
    if (graph) {
        (void)hipGraphDestroy(graph);
        graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    capturing = true;
    int rc = launch_chunk(0, graph_steps);
    capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(NH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    graph = g;
    HIPCHK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    graph_dirty = false;
    return NH_OK;
}

int Arena::alloc_stim(uint32_t cap) {
    stim_cap = cap;
    if (!stim_cols) return NH_OK;
    const size_t ng = std::max<size_t>(stim_group.size(), 1);
    TRY(alloc(d_stim_win, (size_t)stim_cap * stim_cols * 4));
    TRY(alloc(d_stim_nz, (size_t)stim_cap * ng * 4));
    for (int i = 0; i < 2; ++i) {
        // window rows, then the group flags of every row
        HIPCHK(hipHostMalloc((void **)&h_stage[i], ((size_t)stim_cap * stim_cols + stim_cap * ng) * 4, 0));
        HIPCHK(hipEventCreateWithFlags(&stage_ev[i], hipEventDisableTiming));
    }
    return NH_OK;
}

int Arena::upload_stim(uint64_t t_arena, uint32_t k) {
    if (!stim_cols || !packer_) return NH_OK;
    float *h = h_stage[stage_i];
    if (stage_used[stage_i]) HIPCHK(hipEventSynchronize(stage_ev[stage_i]));
    stage_used[stage_i] = true;
    packer_(packer_ctx_, t_arena, k, h);
    // per step and group: does any input differ from +-0?  (a step whose
    // inputs are all zero adds only +-0 to I, which the kernel may skip)
    const size_t ng = stim_group.size();
    uint32_t *nz = (uint32_t *)(h + (size_t)stim_cap * stim_cols);
    for (uint32_t j = 0; j < k; ++j) {
        const float *row = h + (size_t)j * stim_cols;
        for (size_t g = 0; g < ng; ++g) {
            uint32_t any = 0;
            for (uint32_t q = 0; q < stim_group[g].y; ++q) any |= row[stim_group[g].x + q] != 0.0f ? 1u : 0u;
            nz[(size_t)j * ng + g] = any;
        }
    }
    HIPCHK(hipMemcpyAsync(d_stim_win.p, h, (size_t)k * stim_cols * 4, hipMemcpyHostToDevice, stream));
    if (ng) HIPCHK(hipMemcpyAsync(d_stim_nz.p, nz, (size_t)k * ng * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipEventRecord(stage_ev[stage_i], stream));
    stage_i ^= 1;
    return NH_OK;
}

int Arena::run(uint64_t n_steps) {
    if (n_total == 0) {
        t += n_steps;
        last_run_ms = 0.0;
        timing_valid = false;
        return NH_OK;
    }
    TRY(sync_dev_args());
    if (use_graph) {
        // graph chunk: a multiple of D near the requested length (and even
        // in the paired shape, so that every launch of a chunk is a pair)
        const uint32_t want = std::max<uint32_t>(graph_steps_req, 1);
        const uint32_t unit = (D % 2 && pair_usable()) ? 2 * D : D;
        const uint32_t g = std::max<uint32_t>(1, (want + unit / 2) / unit) * unit;
        if (g != graph_steps) {
            graph_steps = g;
            graph_dirty = true;
        }
        if (graph_steps > stim_cap) use_graph = 0;
    }
    if (run_usable()) {
        // persistent kernel: one launch per chunk of steps, no graph
        if (n_steps == 0) return NH_OK;
        TRY(begin_timing());
        while (n_steps) {
            uint64_t k = std::min<uint64_t>(n_steps, run_max_steps);
            if (rast_steps) {
                const uint64_t room = rast_steps - (t - rast_step0);
                if (room == 0) {
                    TRY(drain_raster());
                    continue;
                }
                k = std::min<uint64_t>(k, room);
            }
            TRY(launch_persistent(t, (uint32_t)k));
            pair_launched = true;
            persist_launched = true;
            t += k;
            n_steps -= k;
        }
        return end_timing();
    }
    if (use_graph && graph_dirty) TRY(capture_graph());
    if (n_steps == 0) return NH_OK;  // a zero-step run only prepares the graph
    TRY(begin_timing());
    while (n_steps) {
        uint64_t k = std::min<uint64_t>(n_steps, stim_cap);
        // A captured chunk replays from steps t = 0 (mod D).  A run that
        // reaches a whole chunk after the next such step realigns with a
        // direct piece; a shorter one goes direct in one piece (so a short
        // run in the paired shape is all pairs, whatever its start step).
        const bool graph_ok = use_graph != 0 && t % D == 0;
        if (use_graph && !graph_ok && n_steps >= (D - t % D) + graph_steps)
            k = std::min<uint64_t>(k, D - t % D);
        if (graph_ok) k = std::min<uint64_t>(k, graph_steps);
        if (rast_steps) {
            const uint64_t room = rast_steps - (t - rast_step0);
            if (room == 0) {
                TRY(drain_raster());
                continue;
            }
            k = std::min<uint64_t>(k, room);
        }
        TRY(upload_stim(t, (uint32_t)k));
        if (graph_ok && k == graph_steps) {
            TRY(sync_t_dev(t));
            HIPCHK(hipGraphLaunch(graph_exec, stream));
            t_dev_val += k;
            // a replay counts as its step launches (pairs: k / 2)
            TRY(note_launches(pair_usable() ? k / 2 + k % 2 : k));
        } else {
            TRY(launch_chunk(t, (uint32_t)k));
        }
        if (k > 1 && pair_usable()) pair_launched = true;
        t += k;
        n_steps -= k;
    }
    return end_timing();
}

int Arena::sync_dev_args() {
    if (!d_args.p) {
        TRY(alloc(d_args, sizeof(StepArgs)));
        args_dev_valid = false;
    }
    StepArgs a = args();
    a.k = 0;
    a.slot = 0;
    if (args_dev_valid && std::memcmp(&a, &args_dev_, sizeof(StepArgs)) == 0) return NH_OK;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(d_args.p, &a, sizeof(StepArgs), hipMemcpyHostToDevice));
    args_dev_ = a;
    args_dev_valid = true;
    return NH_OK;
}

// Events of a run: ev0 before its first launch, ev_tail behind its first
// step launch (note_launches), ev1 after its last.
int Arena::begin_timing() {
    timing_valid = false;
    if (!ev_tail) HIPCHK(hipEventCreateWithFlags(&ev_tail, hipEventDisableSystemFence));
    HIPCHK(hipEventRecord(ev0, stream));
    tail_marking = true;
    tail_marked = false;
    tail_launches = 0;
    return NH_OK;
}

int Arena::end_timing() {
    tail_marking = false;
    HIPCHK(hipEventRecord(ev1, stream));
    timing_valid = true;
    return NH_OK;
}

int Arena::sync() {
    HIPCHK(hipStreamSynchronize(stream));
    return check_pair_abort();
}

int Arena::sync_timed() {
    TRY(sync());
    if (timing_valid) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess) last_run_ms = ms;
        last_tail_ms = last_run_ms;
        last_tail_launches = 0;
        if (tail_marked && tail_launches && hipEventElapsedTime(&ms, ev_tail, ev1) == hipSuccess) {
            last_tail_ms = ms;
            last_tail_launches = tail_launches;
        }
        timing_valid = false;
    }
    return NH_OK;
}

int Arena::reset() {
    // (a failed paired launch is what a reset recovers from: no abort check)
    HIPCHK(hipStreamSynchronize(stream));
    timing_valid = false;
    pn_marked = false;
    HIPCHK(hipMemcpyAsync(d_x.p, d_x0.p, d_x0.bytes, hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemsetAsync(d_s.p, 0, d_s.bytes, stream));
    HIPCHK(hipMemsetAsync(d_ring.p, 0, d_ring.bytes, stream));
    if (n_x) {
        HIPCHK(hipMemsetAsync(d_xsyn.p, 0, d_xsyn.bytes, stream));
        HIPCHK(hipMemsetAsync(d_xring.p, 0, d_xring.bytes, stream));
    }
    HIPCHK(hipMemsetAsync(d_cnt.p, 0, d_cnt.bytes, stream));
    HIPCHK(hipMemsetAsync(d_totals.p, 0, d_totals.bytes, stream));
    if (d_bits.p) HIPCHK(hipMemsetAsync(d_bits.p, 0, d_bits.bytes, stream));
    HIPCHK(hipMemsetAsync(d_pn.p, 0, d_pn.bytes, stream));
    HIPCHK(hipMemsetAsync(d_last.p, 0xFF, d_last.bytes, stream));
    HIPCHK(hipMemsetAsync(d_net_spikes.p, 0, d_net_spikes.bytes, stream));
    HIPCHK(hipMemsetAsync(d_rast_ctl.p, 0, d_rast_ctl.bytes, stream));
    HIPCHK(hipMemsetAsync(d_t.p, 0, 8, stream));
    t_dev_val = 0;
    if (d_pub3.p) {
        HIPCHK(hipMemsetAsync(d_pub3.p, 0, d_pub3.bytes, stream));
        HIPCHK(hipMemsetAsync(d_gctr.p, 0, d_gctr.bytes, stream));
        HIPCHK(hipMemsetAsync(d_done.p, 0, d_done.bytes, stream));
    }
    if (d_pub.p) {
        HIPCHK(hipMemsetAsync(d_pub.p, 0, d_pub.bytes, stream));
        HIPCHK(hipMemsetAsync(d_pair_abort.p, 0, d_pair_abort.bytes, stream));
        h_pair_abort.clear();
    }
    HIPCHK(hipStreamSynchronize(stream));
    pair_launched = false;
    persist_launched = false;
    t = 0;
    rast_step0 = 0;
    h_rast_cnt.clear();
    h_rast_ids.clear();
    h_rast_step0 = 0;
    return NH_OK;
}

static std::string shape_name(int f) {
    return f == 3 ? "persistent" : (f == 2 ? "paired" : (f ? "fused" : "two-kernel"));
}

int Arena::autotune(uint32_t trial_steps, std::string &report) {
    if (t != 0) return set_error(NH_ERR_ARG, "autotune runs from the initial state (step 0)");
    if (n_total == 0 || dynamic) return NH_OK;
    struct Cand {
        int fused;
        uint32_t block, graph, bins;
        double ms;
    };
    // bin tables for the one-step shape's LDS-binned push (0 = the
    // wave-by-wave push); only where 32-bit bin keys cover the ring
    std::vector<uint32_t> bin_opts = {0u};
    if ((uint64_t)D * n_total < 0xFFFFFFFFull) bin_opts.insert(bin_opts.end(), {1024u, 4096u});
    std::vector<Cand> cands;
    for (int f : {1, 0, 2, 3})
        for (uint32_t b : {64u, 128u, 256u, 512u})
            for (uint32_t g : {64u, 256u})
                for (uint32_t bins : bin_opts) {
                    if (!use_graph && g != 64u) continue;
                    if (f == 3 && g != 64u) continue;   // no graph: one launch per chunk
                    if (f != 1 && bins) continue;       // bins belong to the one-step shape
                    if (bins && b < 256u) continue;     // a bin table per small workgroup: LDS-bound
                    if (f >= 2) {
                        // only where a paired / persistent launch would actually run
                        const int keep_f = fused;
                        const uint32_t keep_b = block;
                        fused = f;
                        block = b;
                        const bool ok = f == 2 ? pair_usable() : run_usable();
                        fused = keep_f;
                        block = keep_b;
                        if (!ok) continue;
                    }
                    cands.push_back({f, b, g, bins, 0.0});
                }
    Cand best{fused, block, graph_steps_req, lds_bins, 1e30};
    report.clear();
    // one candidate: from step 0, a warm-up, then the faster of two timed
    // trials (a single trial can be off by several per cent)
    auto trial = [&](Cand &c) -> int {
        fused = c.fused;
        block = c.block;
        graph_steps_req = c.graph;
        lds_bins = c.bins;
        graph_dirty = true;
        TRY(reset());
        TRY(run(0));                       // capture outside the timed trials
        TRY(run(std::min<uint32_t>(trial_steps, 2 * D)));  // warm caches
        TRY(sync());
        for (int r = 0; r < 2; ++r) {
            TRY(run(trial_steps));
            TRY(sync_timed());
            c.ms = std::min(c.ms, last_run_ms);
        }
        return NH_OK;
    };
    auto line = [&](const Cand &c) {
        return shape_name(c.fused) + std::string(" block ") + std::to_string(c.block) + " graph " +
               std::to_string(c.graph) + (c.bins ? " bins " + std::to_string(c.bins) : std::string()) + ": " +
               std::to_string(c.ms * 1e3 / trial_steps) + " us/step\n";
    };
    for (Cand &c : cands) {
        c.ms = 1e30;
        TRY(trial(c));
        report += line(c);
    }
    // the three fastest once more, head to head
    std::vector<size_t> order(cands.size());
    for (size_t q = 0; q < order.size(); ++q) order[q] = q;
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cands[x].ms < cands[y].ms; });
    for (size_t q = 0; q < order.size() && q < 3; ++q) {
        Cand &c = cands[order[q]];
        TRY(trial(c));
        report += "again " + line(c);
    }
    for (const Cand &c : cands)
        if (c.ms < best.ms) best = c;
    fused = best.fused;
    block = best.block;
    graph_steps_req = best.graph;
    lds_bins = best.bins;
    graph_dirty = true;
    TRY(reset());
    report += "chosen: " + shape_name(best.fused) + " block " + std::to_string(best.block) + " graph " +
              std::to_string(best.graph) + (best.bins ? " bins " + std::to_string(best.bins) : std::string()) +
              "\n";
    return NH_OK;
}

int Arena::read_x(std::vector<float2> &out) {
    TRY(sync());
    out.resize(n_total);
    if (n_total) HIPCHK(hipMemcpy(out.data(), d_x.p, (size_t)n_total * 8, hipMemcpyDeviceToHost));
    return NH_OK;
}

int Arena::read_f32(const DBuf &b, std::vector<float> &out, size_t n) {
    TRY(sync());
    out.resize(n);
    if (n) HIPCHK(hipMemcpy(out.data(), b.p, n * 4, hipMemcpyDeviceToHost));
    return NH_OK;
}

int Arena::read_u32(const DBuf &b, std::vector<uint32_t> &out, size_t n) {
    TRY(sync());
    out.resize(n);
    if (n) HIPCHK(hipMemcpy(out.data(), b.p, n * 4, hipMemcpyDeviceToHost));
    return NH_OK;
}

int Arena::read_u64(const DBuf &b, std::vector<uint64_t> &out, size_t n) {
    TRY(sync());
    out.resize(n);
    if (n) HIPCHK(hipMemcpy(out.data(), b.p, n * 8, hipMemcpyDeviceToHost));
    return NH_OK;
}

int Arena::count_rows(uint32_t base, uint32_t n, uint64_t rows, uint32_t *dev_out, uint32_t *host_out) {
    if (n == 0) return NH_OK;
    if (!d_bits.p || rows == 0 || rows > bits_rows || rows >= t || base % 64u ||
        (uint64_t)base + ((n + 63u) & ~63u) > n_total)
        return set_error(NH_ERR_ARG, "count_rows: window or range outside the record ring");
    launch_count_rows(d_bits.as<uint64_t>(), bits_rows, words, base, n, t, (uint32_t)rows, dev_out, stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host_out, dev_out, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    return sync();
}

int Arena::read_pn(uint32_t base, uint32_t n, uint32_t *host_out) {
    if (n == 0) return NH_OK;
    if ((uint64_t)base + n > n_total) return set_error(NH_ERR_ARG, "read_pn: range outside the arena");
    HIPCHK(hipMemcpyAsync(host_out, d_pn.as<uint32_t>() + base, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    return sync();
}

// Totals follow from the per-neuron lifetime counts: spikes = sum pn,
// events = sum pn[i] * out-degree(i) (every spike delivers its whole row).
int Arena::spikes_total(uint64_t &out) {
    out = 0;
    if (t == 0 || n_total == 0) return NH_OK;
    std::vector<uint32_t> pn;
    TRY(read_u32(d_pn, pn, n_total));
    for (uint32_t v : pn) out += v;
    return NH_OK;
}

int Arena::spikes_in(uint32_t base, uint32_t n, uint64_t &out) {
    out = 0;
    if (n == 0) return NH_OK;
    TRY(sync());
    std::vector<uint32_t> pn(n);
    HIPCHK(hipMemcpy(pn.data(), d_pn.as<uint32_t>() + base, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (uint32_t v : pn) out += v;
    return NH_OK;
}

int Arena::events_total(uint64_t &out) {
    out = 0;
    if (t == 0 || n_total == 0) return NH_OK;
    std::vector<uint32_t> pn;
    TRY(read_u32(d_pn, pn, n_total));
    for (uint32_t i = 0; i < n_total; ++i)
        if (pn[i]) out += (uint64_t)pn[i] * h_deg_[i];
    return NH_OK;
}

int Arena::mark_counters() {
    if (n_total == 0) {
        pn_marked = true;
        return NH_OK;
    }
    if (!d_pn_mark.p) TRY(alloc(d_pn_mark, (size_t)n_total * 4));  // (an arena built elsewhere)
    HIPCHK(hipMemcpyAsync(d_pn_mark.p, d_pn.p, (size_t)n_total * 4, hipMemcpyDeviceToDevice, stream));
    pn_marked = true;
    return NH_OK;
}

int Arena::counts_since_mark(uint64_t &spikes, uint64_t &events) {
    spikes = events = 0;
    if (!pn_marked) return set_error(NH_ERR_ARG, "no counter mark since creation or reset");
    if (n_total == 0) return NH_OK;
    std::vector<uint32_t> pn, mk;
    TRY(read_u32(d_pn, pn, n_total));
    TRY(read_u32(d_pn_mark, mk, n_total));
    for (uint32_t i = 0; i < n_total; ++i) {
        const uint32_t d = pn[i] - mk[i];
        spikes += d;
        events += (uint64_t)d * h_deg_[i];
    }
    return NH_OK;
}

int Arena::drain_raster() {
    if (!rast_steps) return NH_OK;
    std::vector<uint64_t> ctl;
    TRY(read_u64(d_rast_ctl, ctl, 4));
    if (ctl[2]) return set_error(NH_ERR_CAPACITY, "raster window overflow");
    const uint64_t steps = t - rast_step0;
    if (steps > rast_steps) return set_error(NH_ERR_CAPACITY, "raster window overflow");
    if (steps) {
        std::vector<uint64_t> rows;
        TRY(read_u64(d_rast_bits, rows, (size_t)steps * words));
        if (h_rast_cnt.empty()) h_rast_step0 = rast_step0;
        for (uint64_t s = 0; s < steps; ++s) {
            const uint64_t *row = rows.data() + s * words;
            uint32_t c = 0;
            for (uint32_t w = 0; w < words; ++w) {
                uint64_t m = row[w];
                while (m) {
                    h_rast_ids.push_back(w * 64u + (uint32_t)__builtin_ctzll(m));
                    m &= m - 1;
                    ++c;
                }
            }
            h_rast_cnt.push_back(c);
        }
    }
    rast_step0 = t;
    const uint64_t nc[3] = {0, rast_step0, 0};
    HIPCHK(hipMemcpy(d_rast_ctl.p, nc, sizeof(nc), hipMemcpyHostToDevice));
    return NH_OK;
}

}  // namespace nh
