// batch_api.cpp -- nh_batch: many independent networks of uneven size packed
// into one arena (offset tables + per-wave network index), stepped by the
// same kernels as the plain path in one launch per step.  Every network
// keeps its own seed, noise, stimulus program and raster, so its result is
// the one the plain path gives for it alone (int ring: bit-exact).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "arena.h"

using nh::Arena;
using nh::set_error;

struct nh_batch {
    Arena ar;
    std::vector<nh_config> cfgs;
    // per network, per input: series played from local step play_start
    std::vector<std::vector<std::vector<float>>> series;
    std::vector<std::vector<uint64_t>> play_start;
    // per network raster (split from the arena raster on drain)
    std::vector<std::vector<uint32_t>> r_cnt, r_ids;
    std::vector<uint64_t> r_step0;
    uint64_t split_upto = 0;  // arena steps already split
};

#define TRY(expr)            \
    do {                     \
        int rc_ = (expr);    \
        if (rc_) return rc_; \
    } while (0)

#define CHECK_B(b) \
    if (!(b)) return set_error(NH_ERR_ARG, "null batch")

static void pack_batch(void *ctx, uint64_t t0, uint32_t k, float *out) {
    const nh_batch *b = (const nh_batch *)ctx;
    const uint32_t cols = b->ar.stim_cols;
    for (uint32_t j = 0; j < k; ++j) {
        float *row = out + (size_t)j * cols;
        for (uint32_t q = 0; q < b->ar.n_nets; ++q) {
            const auto &sl = b->ar.slots[q];
            const uint64_t tl = t0 + j - (uint64_t)b->ar.recs[q].t0;
            for (uint32_t in = 0; in < sl.n_inputs; ++in) {
                const auto &ser = b->series[q][in];
                const uint64_t idx = tl - b->play_start[q][in];
                row[sl.stim_base + in] = (!ser.empty() && idx < ser.size()) ? ser[idx] : 0.0f;
            }
        }
    }
}

// Move drained arena raster rows into per-network rasters (local ids).
static int split_raster(nh_batch *b) {
    auto &ar = b->ar;
    TRY(ar.drain_raster());
    const uint64_t steps = ar.h_rast_cnt.size();
    uint64_t o = 0;
    for (uint64_t s = 0; s < steps; ++s) {
        const uint32_t c = ar.h_rast_cnt[s];
        // ids of one step are ascending and networks are laid out in order,
        // so each network's ids form one contiguous run
        std::vector<uint32_t> per(ar.n_nets, 0);
        for (uint32_t q = 0; q < c; ++q) {
            const uint32_t id = ar.h_rast_ids[o + q];
            // binary search the owning network
            uint32_t lo = 0, hi = ar.n_nets;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (ar.slots[mid].base <= id) lo = mid;
                else hi = mid;
            }
            per[lo]++;
            b->r_ids[lo].push_back(id - ar.slots[lo].base);
        }
        for (uint32_t k = 0; k < ar.n_nets; ++k) b->r_cnt[k].push_back(per[k]);
        o += c;
    }
    ar.h_rast_cnt.clear();
    ar.h_rast_ids.clear();
    return NH_OK;
}

extern "C" {

int nh_batch_create(uint32_t n_nets, const nh_net_desc *nets, const nh_config *cfgs, nh_batch **out) {
    return nh::guarded([&]() -> int {
        if (!nets || !cfgs || !out) return set_error(NH_ERR_ARG, "null argument");
        *out = nullptr;
        for (uint32_t k = 0; k < n_nets; ++k) TRY(nh_check_net(&nets[k], &cfgs[k]));
        int dev = 0;
        if (hipGetDeviceCount(&dev) != hipSuccess || dev < 1) return set_error(NH_ERR_NO_DEVICE, "no HIP device");
        std::unique_ptr<nh_batch> b(new (std::nothrow) nh_batch());
        if (!b) return set_error(NH_ERR_CAPACITY, "out of host memory");
        for (uint32_t k = 1; k < n_nets; ++k)
            if (cfgs[k].dt_ms != cfgs[0].dt_ms) return set_error(NH_ERR_ARG, "packed networks must share dt");
        nh::Tuning tune;
        tune.fused = cfgs[0].step_shape;
        tune.use_graph = cfgs[0].exec_mode == 0 ? -1 : (cfgs[0].exec_mode == 2 ? 1 : 0);
        tune.graph_steps = cfgs[0].graph_steps ? cfgs[0].graph_steps : 64;
        tune.block = cfgs[0].block;
        tune.lds_bins = cfgs[0].lds_bins;
        TRY(b->ar.build(n_nets, nets, cfgs, cfgs[0].raster_steps, tune));
        b->cfgs.assign(cfgs, cfgs + n_nets);
        b->series.resize(n_nets);
        b->play_start.resize(n_nets);
        for (uint32_t k = 0; k < n_nets; ++k) {
            b->series[k].assign(nets[k].n_inputs, {});
            b->play_start[k].assign(nets[k].n_inputs, 0);
        }
        b->r_cnt.resize(n_nets);
        b->r_ids.resize(n_nets);
        b->r_step0.assign(n_nets, 0);
        b->ar.set_stim_packer(pack_batch, b.get());
        *out = b.release();
        return NH_OK;
    });
}

void nh_batch_destroy(nh_batch *b) { delete b; }

int nh_batch_run_steps(nh_batch *b, uint64_t n_steps) {
    return nh::guarded([&]() -> int {
        CHECK_B(b);
        return b->ar.run(n_steps);
    });
}

int nh_batch_sync(nh_batch *b) {
    CHECK_B(b);
    return b->ar.sync();
}

int nh_batch_last_run_ms(nh_batch *b, double *ms) {
    CHECK_B(b);
    TRY(b->ar.sync_timed());
    *ms = b->ar.last_run_ms;
    return NH_OK;
}

int nh_batch_autotune(nh_batch *b, uint32_t trial_steps, char *report, uint32_t report_cap) {
    CHECK_B(b);
    std::string rep;
    TRY(b->ar.autotune(trial_steps ? trial_steps : 256, rep));
    for (auto &c : b->r_cnt) c.clear();
    for (auto &v : b->r_ids) v.clear();
    std::fill(b->r_step0.begin(), b->r_step0.end(), 0);
    if (report && report_cap) {
        std::strncpy(report, rep.c_str(), report_cap - 1);
        report[report_cap - 1] = 0;
    }
    return NH_OK;
}

int nh_batch_spikes_total(nh_batch *b, uint32_t net, uint64_t *out) {
    CHECK_B(b);
    if (net >= b->ar.n_nets) return set_error(NH_ERR_ARG, "net out of range");
    return b->ar.spikes_in(b->ar.slots[net].base, b->ar.slots[net].n, *out);
}

int nh_batch_events_total(nh_batch *b, uint64_t *out) {
    CHECK_B(b);
    return b->ar.events_total(*out);
}

int nh_batch_mark_counters(nh_batch *b) {
    return nh::guarded([&]() -> int {
        CHECK_B(b);
        return b->ar.mark_counters();
    });
}

int nh_batch_counts_since_mark(nh_batch *b, uint64_t *spikes, uint64_t *events) {
    return nh::guarded([&]() -> int {
        CHECK_B(b);
        if (!spikes || !events) return set_error(NH_ERR_ARG, "null argument");
        return b->ar.counts_since_mark(*spikes, *events);
    });
}

int nh_batch_read_state(nh_batch *b, uint32_t net, uint32_t slot, float *out, uint32_t n) {
    CHECK_B(b);
    if (net >= b->ar.n_nets || slot >= 2) return set_error(NH_ERR_ARG, "net/slot out of range");
    const auto &sl = b->ar.slots[net];
    if (n < sl.n) return set_error(NH_ERR_ARG, "output too small");
    std::vector<float2> x;
    TRY(b->ar.read_x(x));
    for (uint32_t i = 0; i < sl.n; ++i) out[i] = slot == 0 ? x[sl.base + i].x : x[sl.base + i].y;
    return NH_OK;
}

int nh_batch_read_spikes(nh_batch *b, uint32_t net, uint64_t *step0, uint32_t *counts, uint64_t max_steps,
                         uint32_t *ids, uint64_t max_ids, uint64_t *n_steps, uint64_t *n_ids) {
    return nh::guarded([&]() -> int {
        CHECK_B(b);
        if (net >= b->ar.n_nets) return set_error(NH_ERR_ARG, "net out of range");
        if (!b->ar.rast_steps) return set_error(NH_ERR_UNSUPPORTED, "raster stream off (raster_steps = 0)");
        TRY(split_raster(b));
        auto &c = b->r_cnt[net];
        auto &v = b->r_ids[net];
        *n_steps = c.size();
        *n_ids = v.size();
        *step0 = b->r_step0[net];
        if (c.size() > max_steps || v.size() > max_ids)
            return set_error(NH_ERR_CAPACITY, "raster output arrays too small");
        std::copy(c.begin(), c.end(), counts);
        std::copy(v.begin(), v.end(), ids);
        b->r_step0[net] += c.size();
        c.clear();
        v.clear();
        return NH_OK;
    });
}

// Backend::spike_counts (sim/src/lib.rs:119-184; cuda.rs:725-763) of one
// packed network: lifetime counts when rows >= t, the window counted on the
// device when rows <= the record ring, NH_ERR_WINDOW between the two.
int nh_batch_spike_counts(nh_batch *b, uint32_t net, uint64_t rows, uint32_t *out, uint32_t n) {
    return nh::guarded([&]() -> int {
        CHECK_B(b);
        if (net >= b->ar.n_nets) return set_error(NH_ERR_ARG, "net out of range");
        if (!out) return set_error(NH_ERR_ARG, "null output");
        const auto &sl = b->ar.slots[net];
        if (n < sl.n) return set_error(NH_ERR_ARG, "output too small");
        std::memset(out, 0, (size_t)sl.n * 4);
        const uint64_t t = b->ar.t;
        if (!b->cfgs[0].record_spikes || t == 0 || sl.n == 0 || rows == 0) return NH_OK;
        if (rows >= t) return b->ar.read_pn(sl.base, sl.n, out);
        const uint32_t R = b->ar.bits_rows;
        if (rows > R)
            return set_error(NH_ERR_WINDOW, "spike_counts window " + std::to_string(rows) +
                                                " exceeds the recorded " + std::to_string(R) + " steps");
        if (!b->ar.d_counts.p) {
            const hipError_t e = b->ar.d_counts.alloc((size_t)b->ar.n_total * 4);
            if (e != hipSuccess) return set_error(NH_ERR_HIP, std::string("count buffer: ") + hipGetErrorString(e));
            b->ar.dev_bytes += b->ar.d_counts.bytes;
        }
        return b->ar.count_rows(sl.base, sl.n, rows, b->ar.d_counts.as<uint32_t>(), out);
    });
}

int nh_batch_play_input(nh_batch *b, uint32_t net, uint32_t input, const float *series, uint64_t len) {
    CHECK_B(b);
    if (net >= b->ar.n_nets || input >= b->ar.slots[net].n_inputs)
        return set_error(NH_ERR_ARG, "net/input out of range");
    b->series[net][input].assign(series, series + len);
    b->play_start[net][input] = b->ar.t - (uint64_t)b->ar.recs[net].t0;
    return NH_OK;
}

}  // extern "C"
