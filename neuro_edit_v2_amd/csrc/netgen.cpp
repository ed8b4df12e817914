// netgen.cpp -- network generators behind nh_net_* (the reference's
// net-format builders restated in C++, plus the synthetic workloads of the
// benchmark configurations).  Output is the flat simulator layout: CSR by
// source (Dale-signed weights, delays in steps), model table, param and
// initial-state rows, stimulus map.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/neuro_hip.h"
#include "host_util.h"
#include "netgen.h"

using nh::set_error;

namespace {

// xoshiro256** seeded by splitmix64 -- net-format/src/rng.rs:17-81
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t v = z;
            v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
            v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
            s[i] = v ^ (v >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double f64() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    float f32() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    double uniform(double lo, double hi) { return lo + (hi - lo) * f64(); }
    uint64_t below(uint64_t n) { return next() % (n ? n : 1); }
};

const float kRS[4] = {0.02f, 0.2f, -65.0f, 8.0f};   // izh.rs:140
const float kFS[4] = {0.1f, 0.2f, -65.0f, 2.0f};    // izh.rs:146
// adex.rs:22-26 (TONIC, NONADAP)
const float kTonic[10] = {281.0f, 30.0f, -70.6f, -50.4f, 2.0f, 4.0f, 144.0f, 80.5f, -70.6f, 20.0f};
const float kNonAdap[10] = {281.0f, 30.0f, -70.6f, -50.4f, 2.0f, 0.0f, 144.0f, 0.0f, -70.6f, 20.0f};

}  // namespace


// Per-neuron rows for a one- or two-kind net: izh init x = [-65, b*-65]
// (izh.rs:300-303), adex x = [EL, 0] (adex.rs:149-152).
static void set_rows(nh_net &net, const std::vector<uint8_t> &inh, int model) {
    const uint32_t n = net.n;
    net.model_ids = {(uint8_t)model};
    net.model.assign(n, 0);
    net.n_params = model == 1 ? 10 : 4;
    net.n_state = 2;
    net.p.assign((size_t)n * net.n_params, 0.0f);
    net.x0.assign((size_t)n * 2, 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        const float *row = model == 1 ? (inh[i] ? kNonAdap : kTonic) : (inh[i] ? kFS : kRS);
        std::memcpy(&net.p[(size_t)i * net.n_params], row, sizeof(float) * net.n_params);
        if (model == 1) {
            net.x0[2 * (size_t)i] = row[2];
            net.x0[2 * (size_t)i + 1] = 0.0f;
        } else {
            net.x0[2 * (size_t)i] = -65.0f;
            net.x0[2 * (size_t)i + 1] = row[1] * -65.0f;
        }
    }
}

extern "C" {

void nh_net_free(nh_net *net) { delete net; }

int nh_net_view(const nh_net *net, nh_net_desc *d) {
    if (!net || !d) return set_error(NH_ERR_ARG, "null argument");
    std::memset(d, 0, sizeof(*d));
    d->n = net->n;
    d->n_syn = net->tgt.size();
    d->src_ptr = net->src_ptr.data();
    d->tgt = net->tgt.data();
    d->w = net->w.data();
    d->delay = net->delay.data();
    d->tau = net->tau.empty() ? nullptr : net->tau.data();
    d->n_models = (uint32_t)net->model_ids.size();
    d->model_ids = net->model_ids.data();
    d->model = net->model.data();
    d->n_params = net->n_params;
    d->n_state = net->n_state;
    d->p = net->p.data();
    d->x0 = net->x0.data();
    d->n_inputs = net->n_inputs;
    d->stim_ptr = net->n_inputs ? net->stim_ptr.data() : nullptr;
    d->stim_input = net->stim_input.data();
    d->stim_w = net->stim_w.data();
    return NH_OK;
}

// bench::bench_net -- sim/src/bench.rs:10-53, then synapse_csr
// (net-format/src/network.rs:654-692) at dt.
int nh_net_bench(uint32_t n, uint32_t k, uint64_t seed, float w_scale, float w_inh, float dt_ms,
                 nh_net **out) {
    return nh::guarded([&]() -> int {
        if (!out) return set_error(NH_ERR_ARG, "null out");
        std::unique_ptr<nh_net> net(new (std::nothrow) nh_net());
        if (!net) return set_error(NH_ERR_CAPACITY, "out of host memory");
        Rng rng(seed);
        const uint64_t n_exc = (uint64_t)n * 4 / 5;
        for (uint32_t i = 0; i < n; ++i) {
            (void)rng.uniform(-500.0, 500.0);  // positions: drawn for stream parity, unused by
            (void)rng.uniform(-500.0, 500.0);  // the bench (delays are set directly)
            (void)rng.uniform(-500.0, 500.0);
        }
        net->n = n;
        net->src_ptr.assign(n + 1, 0);
        net->tgt.reserve((size_t)n * k);
        net->w.reserve((size_t)n * k);
        net->delay.reserve((size_t)n * k);
        for (uint32_t src = 0; src < n; ++src) {
            const bool exc = src < n_exc;
            const float wm = exc ? 6.0f * w_scale : w_inh;
            for (uint32_t q = 0; q < k; ++q) {
                const uint32_t tgt = (uint32_t)rng.below(n);
                const float delay_ms = exc ? (float)rng.uniform(1.0, 21.0) : 1.0f;
                float steps = std::round(delay_ms / dt_ms);
                if (steps < 1.0f) steps = 1.0f;
                if (!std::isfinite(steps) || steps > 65535.0f) return set_error(NH_ERR_NET, "delay exceeds the 65535-step ring");
                net->tgt.push_back(tgt);
                net->w.push_back(exc ? std::fabs(wm) : -std::fabs(wm));
                net->delay.push_back((uint16_t)steps);
            }
            net->src_ptr[src + 1] = (uint32_t)net->tgt.size();
        }
        std::vector<uint8_t> inh(n);
        for (uint32_t i = 0; i < n; ++i) inh[i] = i >= n_exc;
        set_rows(*net, inh, 0);
        *out = net.release();
        return NH_OK;
    });
}

// Random-cube workload: positions U[-500,500)^3 (bench.rs:17-21 order),
// first 80% RS, last 20% FS; per source an out-degree ~ Poisson(mean) drawn
// by inversion (sequential search from k = 0 with pmf recursion in f64);
// targets uniform without self-connections; w = U[0,1)*w_scale magnitude,
// negative from inhibitory sources (Dale); delay in steps
// clamp(1 + floor(19 d / dmax), 1, 20), d the f64 distance of the f32
// positions, dmax = 1000*sqrt(3).
int nh_net_cube(uint32_t n, double mean_deg, uint64_t seed, float w_scale, nh_net **out) {
    return nh::guarded([&]() -> int {
        if (!out) return set_error(NH_ERR_ARG, "null out");
        if (n > (1u << 27)) return set_error(NH_ERR_ARG, "n too large");
        std::unique_ptr<nh_net> net(new (std::nothrow) nh_net());
        if (!net) return set_error(NH_ERR_CAPACITY, "out of host memory");
        Rng rng(seed);
        std::vector<float> pos((size_t)n * 3);
        for (size_t i = 0; i < (size_t)n * 3; ++i) pos[i] = (float)rng.uniform(-500.0, 500.0);
        const uint64_t n_exc = (uint64_t)n * 4 / 5;
        const double p0 = std::exp(-mean_deg);
        const double dmax = 1000.0 * std::sqrt(3.0);
        net->n = n;
        net->src_ptr.assign(n + 1, 0);
        const size_t guess = (size_t)((double)n * mean_deg * 1.01) + 16;
        net->tgt.reserve(guess);
        net->w.reserve(guess);
        net->delay.reserve(guess);
        for (uint32_t s = 0; s < n; ++s) {
            const double u = rng.f64();
            uint32_t deg = 0;
            double pk = p0, F = p0;
            while (u > F && deg < 100000) {
                ++deg;
                pk *= mean_deg / (double)deg;
                F += pk;
            }
            const bool inh = s >= n_exc;
            for (uint32_t q = 0; q < deg; ++q) {
                uint32_t t;
                do {
                    t = (uint32_t)rng.below(n);
                } while (t == s && n > 1);
                const float wm = rng.f32() * w_scale;
                const double dx = (double)pos[3 * (size_t)s] - pos[3 * (size_t)t];
                const double dy = (double)pos[3 * (size_t)s + 1] - pos[3 * (size_t)t + 1];
                const double dz = (double)pos[3 * (size_t)s + 2] - pos[3 * (size_t)t + 2];
                const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                int dl = 1 + (int)std::floor(19.0 * d / dmax);
                dl = std::min(std::max(dl, 1), 20);
                net->tgt.push_back(t);
                net->w.push_back(inh ? -std::fabs(wm) : std::fabs(wm));
                net->delay.push_back((uint16_t)dl);
            }
            net->src_ptr[s + 1] = (uint32_t)net->tgt.size();
        }
        std::vector<uint8_t> inhv(n);
        for (uint32_t i = 0; i < n; ++i) inhv[i] = i >= n_exc;
        set_rows(*net, inhv, 0);
        *out = net.release();
        return NH_OK;
    });
}

// 3D-grid liquid.  Lattice as builders::create_3d_grid (builders.rs:88-131)
// at spacing 60 (the LSM examples' LIQ_SPACING) centred on the origin;
// kinds by set_exc_inh (builders.rs:262-277, one f64 draw per member);
// wiring p = scale * exp(-(d/(lambda*60))^2) as create_connections
// (builders.rs:150-181) but only over lattice offsets within 3*lambda
// (beyond it p < scale*e^-9), with p precomputed per offset -- the draw
// order differs from the reference's all-pairs loop; delays from the
// conduction law d/100 ms (network.rs:596-601) rounded at dt
// (network.rs:674-685).  Inputs as examples/lsm_deep.rhai:82-95: pairs
// (0,1),(2,3),.. share a pool of `shared` random targets, plus `priv`
// private ones per input, weight w_in.  AdEx nets scale every weight by C
// (281 pF) so one unit moves v as it does for izh (the noise-gain rule of
// sim/src/model/adex.rs:187-191).
int nh_net_grid(uint32_t ex, uint32_t ey, uint32_t ez, double scale, double lambda, double inh_frac,
                float w_exc, float w_inh, int model, uint32_t n_inputs, uint32_t per_input, float w_in,
                uint64_t seed, float dt_ms, nh_net **out) {
    return nh::guarded([&]() -> int {
        if (!out) return set_error(NH_ERR_ARG, "null out");
        const uint64_t n64 = (uint64_t)ex * ey * ez;
        if (n64 == 0 || n64 > (1u << 26)) return set_error(NH_ERR_ARG, "bad grid size");
        std::unique_ptr<nh_net> net(new (std::nothrow) nh_net());
        if (!net) return set_error(NH_ERR_CAPACITY, "out of host memory");
        const uint32_t n = (uint32_t)n64;
        const float spacing = 60.0f, speed = 100.0f;
        Rng rng(seed);
        // lattice in create_3d_grid order: z outer, x, y inner
        std::vector<uint32_t> gx(n), gy(n), gz(n);
        std::vector<float> pos((size_t)n * 3);
        {
            uint32_t id = 0;
            for (uint32_t z = 0; z < ez; ++z)
                for (uint32_t x = 0; x < ex; ++x)
                    for (uint32_t y = 0; y < ey; ++y) {
                        gx[id] = x;
                        gy[id] = y;
                        gz[id] = z;
                        pos[3 * (size_t)id] = ((float)x - (float)ex / 2.0f) * spacing;
                        pos[3 * (size_t)id + 1] = ((float)y - (float)ey / 2.0f) * spacing;
                        pos[3 * (size_t)id + 2] = ((float)z - (float)ez / 2.0f) * spacing;
                        ++id;
                    }
        }
        auto id_of = [&](uint32_t x, uint32_t y, uint32_t z) { return (z * ex + x) * ey + y; };
        std::vector<uint8_t> inh(n);
        for (uint32_t i = 0; i < n; ++i) inh[i] = rng.f64() < inh_frac ? 1 : 0;
        // offsets within the cutoff, with their probability and delay
        const double lam = lambda * spacing;
        const int R = (int)std::ceil(3.0 * lambda);
        struct Off {
            int dx, dy, dz;
            double p;
            uint16_t delay;
        };
        std::vector<Off> offs;
        for (int dz = -R; dz <= R; ++dz)
            for (int dx = -R; dx <= R; ++dx)
                for (int dy = -R; dy <= R; ++dy) {
                    if (!dx && !dy && !dz) continue;
                    const float fx = (float)dx * spacing, fy = (float)dy * spacing, fz = (float)dz * spacing;
                    const float df = std::sqrt(fx * fx + fy * fy + fz * fz);
                    const double d = (double)df;
                    if (d > 3.0 * lam) continue;
                    const double q = d / lam;
                    const double p = scale * std::exp(-(q * q));
                    float dms = df / speed;
                    if (dms < 1e-3f) dms = 1e-3f;
                    float steps = std::round(dms / dt_ms);
                    if (steps < 1.0f) steps = 1.0f;
                    if (steps > 65535.0f) return set_error(NH_ERR_NET, "delay exceeds the 65535-step ring");
                    offs.push_back({dx, dy, dz, p, (uint16_t)steps});
                }
        const float wmul = model == 1 ? kTonic[0] : 1.0f;
        net->n = n;
        net->src_ptr.assign(n + 1, 0);
        for (uint32_t s = 0; s < n; ++s) {
            for (const Off &o : offs) {
                const int x = (int)gx[s] + o.dx, y = (int)gy[s] + o.dy, z = (int)gz[s] + o.dz;
                if (x < 0 || y < 0 || z < 0 || x >= (int)ex || y >= (int)ey || z >= (int)ez) continue;
                if (rng.f64() < o.p) {
                    const float wm = (inh[s] ? w_inh : w_exc) * wmul;
                    net->tgt.push_back(id_of((uint32_t)x, (uint32_t)y, (uint32_t)z));
                    net->w.push_back(inh[s] ? -std::fabs(wm) : std::fabs(wm));
                    net->delay.push_back(o.delay);
                }
            }
            net->src_ptr[s + 1] = (uint32_t)net->tgt.size();
        }
        set_rows(*net, inh, model);
        // inputs
        net->n_inputs = n_inputs;
        if (n_inputs) {
            std::vector<std::vector<uint32_t>> tg(n_inputs);
            const uint32_t shared = per_input * 2;
            for (uint32_t p0 = 0; p0 < n_inputs; p0 += 2) {
                std::vector<uint32_t> pool(shared);
                for (auto &m : pool) m = (uint32_t)rng.below(n);
                for (uint32_t sidx = p0; sidx < std::min(p0 + 2, n_inputs); ++sidx) {
                    tg[sidx] = pool;
                    for (uint32_t q = 0; q < per_input; ++q) tg[sidx].push_back((uint32_t)rng.below(n));
                }
            }
            // StimMap::from_network (sim/src/stim.rs:34-57)
            std::vector<std::vector<std::pair<uint32_t, float>>> per(n);
            for (uint32_t k = 0; k < n_inputs; ++k)
                for (uint32_t m : tg[k]) per[m].push_back({k, w_in * wmul});
            net->stim_ptr.assign(n + 1, 0);
            for (uint32_t i = 0; i < n; ++i) {
                net->stim_ptr[i + 1] = net->stim_ptr[i] + (uint32_t)per[i].size();
                for (auto &e : per[i]) {
                    net->stim_input.push_back(e.first);
                    net->stim_w.push_back(e.second);
                }
            }
        }
        *out = net.release();
        return NH_OK;
    });
}

struct nh_rng_stream {
    Rng rng;
    uint32_t n_src;
    float p;
    nh_rng_stream(uint64_t seed, uint32_t n, float pp) : rng(seed), n_src(n), p(pp) {}
};

int nh_rng_stream_new(uint64_t seed, uint32_t n_src, float p, nh_rng_stream **out) {
    return nh::guarded([&]() -> int {
        if (!out) return set_error(NH_ERR_ARG, "null out");
        *out = new (std::nothrow) nh_rng_stream(seed, n_src, p);
        return *out ? NH_OK : set_error(NH_ERR_CAPACITY, "out of host memory");
    });
}

void nh_rng_stream_free(nh_rng_stream *s) { delete s; }

// RngStream::window -- sim/src/script.rs:2867-2894
int nh_rng_stream_window(nh_rng_stream *s, uint32_t steps, int event, const uint32_t *pairs,
                         uint32_t n_pairs, float strength, float *out) {
    return nh::guarded([&]() -> int {
        if (!s || (!out && steps)) return set_error(NH_ERR_ARG, "null argument");
        const double p = (double)s->p;
        for (uint32_t t = 0; t < steps; ++t) {
            float *row = out + (size_t)t * s->n_src;
            for (uint32_t k = 0; k < s->n_src; ++k) row[k] = s->rng.f64() < p ? 1.0f : 0.0f;
            if (event && strength > 0.0f) {
                for (uint32_t q = 0; q < n_pairs; ++q) {
                    const uint32_t i = pairs[2 * q], j = pairs[2 * q + 1];
                    if (i >= s->n_src || j >= s->n_src) continue;
                    if (s->rng.f64() < (double)strength) {
                        const float bit = s->rng.f64() < p ? 1.0f : 0.0f;
                        row[i] = bit;
                        row[j] = bit;
                    }
                }
            }
        }
        return NH_OK;
    });
}

}  // extern "C"
