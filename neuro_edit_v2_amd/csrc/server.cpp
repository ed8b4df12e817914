// server.cpp -- the served form: a request loop over one resident arena.
//
// The reference server runs each job on its own thread with its own
// simulator and CUDA context (server/src/jobs.rs:106-161,226-339); GPU work
// of concurrent jobs only meets in the driver.  Here one worker thread owns
// one arena sized from free HBM.  Requests are admitted while neurons,
// synapses, stimulus entries and slots are free (first fit over 64-neuron
// units); an admitted request's network, state and delay-ring columns stay
// resident until it finishes; every scheduling round advances all in-flight
// requests together through the same captured step graph.  Each request
// sees only its own local step (noise hash, stimulus, windows), so its
// result equals the plain path's result for it alone.
#include <algorithm>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "arena.h"

namespace nh {
void launch_clear_range(const StepArgs &a, uint32_t base, uint32_t len, hipStream_t st);
}

using nh::Arena;
using nh::set_error;

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// First-fit allocator over [0, cap) with coalescing frees.
struct FreeList {
    std::map<uint64_t, uint64_t> free;  // start -> length
    void init(uint64_t cap) {
        free.clear();
        if (cap) free[0] = cap;
    }
    bool alloc(uint64_t len, uint64_t &at) {
        if (len == 0) {
            at = 0;
            return true;
        }
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second >= len) {
                at = it->first;
                const uint64_t rest = it->second - len;
                free.erase(it);
                if (rest) free[at + len] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t at, uint64_t len) {
        if (len == 0) return;
        auto it = free.emplace(at, len).first;
        auto nx = std::next(it);
        if (nx != free.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free.erase(nx);
        }
        if (it != free.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free.erase(it);
            }
        }
    }
};

struct Request {
    uint64_t id = 0;
    uint32_t n = 0, n_al = 0;
    uint64_t E = 0, n_stim = 0;
    nh_config cfg{};
    uint64_t total_steps = 0;
    uint32_t win_steps = 0;
    // prepared, request-local
    std::vector<uint32_t> row_off;    // n+1 into syn
    std::vector<uint32_t> syn;        // local tgt | delay << 27
    std::vector<float> wf;
    std::vector<int32_t> wq;
    std::vector<nh::PRow> rows;       // distinct parameter rows
    std::vector<uint8_t> row_of;      // per neuron
    std::vector<float2> x0;
    std::vector<uint32_t> stim_off;   // n+1
    std::vector<uint32_t> stim_in;
    std::vector<float> stim_w;
    std::vector<float> stim;          // [steps][n_inputs]
    uint64_t stim_steps = 0;
    // delay-1 in-table (two steps per launch), request-local: per wave the
    // entry count, entries local source | target lane << 26, weights
    std::vector<uint32_t> pl_cnt;
    std::vector<uint32_t> pl_ent;
    std::vector<uint32_t> pl_w;       // f32 or i32 bits, as the ring
    uint64_t E1 = 0;
    // runtime
    uint32_t slot = 0;
    uint64_t base = 0, sbase = 0, stbase = 0, plbase = 0, t0 = 0;
    uint64_t submit_ns = 0, admit_ns = 0, done_ns = 0, spikes = 0;
    bool done = false;
    int error = 0;
    std::string error_msg;
    std::vector<uint32_t> counts;     // [n_windows][n]
};

// Pinned host staging for admission uploads: every array of a request is
// packed here and copied with asynchronous copies (pageable sources would
// make each copy a synchronous staged one).  Reused; grown on demand; free
// again once the admission's stream synchronisation returns.
struct Stage {
    uint8_t *buf = nullptr;
    size_t cap = 0, used = 0;
    Stage() = default;
    Stage(const Stage &) = delete;
    Stage &operator=(const Stage &) = delete;
    ~Stage() {
        if (buf) (void)hipHostFree(buf);
    }
    hipError_t reserve(size_t bytes) {
        used = 0;
        if (bytes <= cap) return hipSuccess;
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 2;
        hipError_t e = hipHostMalloc((void **)&buf, want, 0);
        if (e == hipSuccess) cap = want;
        return e;
    }
    template <class T> T *take(uint64_t count) {
        T *p = (T *)(buf + used);
        used += (count * sizeof(T) + 255) & ~size_t(255);
        assert(used <= cap);
        return p;
    }
};

}  // namespace

struct nh_server {
    Arena ar;
    nh_server_config cfg{};
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<Request *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    std::vector<Request *> slot_req;
    FreeList fn, fs, fst;             // neurons (64-units), synapses, stimulus entries
    FreeList fpl;                     // delay-1 in-table entries (paired shape)
    bool pairs = false;               // two steps per launch where the grid allows
    Stage stage;                      // pinned staging of admission uploads
    std::vector<nh::PRow> table;
    nh_server_stats st{};
    uint64_t active = 0;
    int worker_error = 0;
    std::string worker_msg;
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

namespace {

template <class T> int put(nh::DBuf &b, uint64_t off, const T *src, uint64_t count, hipStream_t st) {
    if (!count) return NH_OK;
    HIPCHK(hipMemcpyAsync((T *)b.p + off, src, count * sizeof(T), hipMemcpyHostToDevice, st));
    return NH_OK;
}

void update_grid(nh_server *s);

// Rows [i0, i1) of a request's CSR into its device layout: each row sorted
// by (delay, target), ties in input order, the local target packed with the
// delay (key = packed word << 32 | input index, so integer order is that
// order), weights quantised for the int ring.  Every synapse is kept, so
// row i lands at src_ptr[i] and ranges of rows are independent.  Returns
// an error message or null.
const char *prep_rows(const nh_net_desc *d, const nh_config *cfg, uint32_t max_delay, bool int_w,
                      uint32_t i0, uint32_t i1, Request *r) {
    std::vector<uint64_t> key;
    for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t lo = d->src_ptr[i], hi = d->src_ptr[i + 1], m = hi - lo;
        key.resize(m);
        for (uint32_t j = lo; j < hi; ++j) {
            if (d->tgt[j] >= d->n) return "synapse target out of range";
            if (d->delay[j] < 1 || d->delay[j] > max_delay) return "delay outside 1..max_delay of the server";
            if (d->tau && d->tau[j] > 0.0f) return "custom-tau synapses are not supported";
            key[j - lo] = ((uint64_t)(d->tgt[j] | ((uint32_t)d->delay[j] << 27)) << 32) | j;
        }
        if (m > 32) {
            std::sort(key.begin(), key.end());
        } else {
            for (uint32_t a = 1; a < m; ++a) {  // short rows: insertion sort
                const uint64_t v = key[a];
                uint32_t b = a;
                for (; b && key[b - 1] > v; --b) key[b] = key[b - 1];
                key[b] = v;
            }
        }
        for (uint32_t q = 0; q < m; ++q) {
            const uint32_t j = (uint32_t)key[q];
            r->syn[lo + q] = (uint32_t)(key[q] >> 32);
            if (int_w) r->wq[lo + q] = (int32_t)std::lrintf(d->w[j] * cfg->int_scale);
            else r->wf[lo + q] = d->w[j];
        }
    }
    return nullptr;
}

// Stimulus window of every slot: local step (t - t0) of each active request.
void pack_server(void *ctx, uint64_t t0, uint32_t k, float *out) {
    nh_server *s = (nh_server *)ctx;
    const uint32_t ni = s->cfg.n_inputs, cols = s->ar.stim_cols;
    std::memset(out, 0, (size_t)k * cols * 4);
    for (uint32_t q = 0; q < s->cfg.max_requests; ++q) {
        Request *r = s->slot_req[q];
        if (!r || !ni) continue;
        for (uint32_t j = 0; j < k; ++j) {
            const uint64_t tl = t0 + j - r->t0;
            if (tl >= r->stim_steps) continue;
            std::memcpy(out + (size_t)j * cols + (size_t)q * ni, &r->stim[tl * ni], ni * 4);
        }
    }
}

int build_arena(nh_server *s) {
    Arena &ar = s->ar;
    const nh_server_config &c = s->cfg;
    ar.dynamic = true;
    ar.n_nets = c.max_requests;
    ar.n_total = (uint32_t)c.cap_neurons;
    ar.words = ar.n_total / 64;
    // two steps per launch needs the delay-1 in-tables (sources packed in
    // 26 bits) and one ring slot more than the deepest delay
    s->pairs = c.cap_neurons < (1ull << 26);
    ar.D = c.max_delay + (s->pairs ? 2 : 1);
    ar.ring_int = c.int_weights ? 1 : 0;
    ar.pmode = 2;
    ar.mmode = 2;
    ar.packed = 1;
    ar.fused = 1;
    ar.stim_cols = c.max_requests * c.n_inputs;
    ar.P_ = nh::kMaxParams;
    ar.n_win = c.n_windows;
    ar.recs.assign(c.max_requests, nh::NetRec{});
    for (auto &r : ar.recs) {
        std::memset(&r, 0, sizeof(r));
        r.active = 0;
    }
    ar.slots.resize(c.max_requests);
    HIPCHK(hipGetDevice(&ar.device));
    HIPCHK(hipStreamCreateWithFlags(&ar.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ar.ev0, hipEventDisableSystemFence));
    HIPCHK(hipEventCreateWithFlags(&ar.ev1, hipEventDisableSystemFence));
    const uint64_t n = ar.n_total;
    auto A = [&](nh::DBuf &b, uint64_t bytes) -> int {
        HIPCHK(b.alloc(bytes));
        ar.dev_bytes += b.bytes;
        return NH_OK;
    };
    TRY(A(ar.d_t, 8));
    TRY(A(ar.d_nets, (uint64_t)c.max_requests * sizeof(nh::NetRec)));
    HIPCHK(hipMemcpy(ar.d_nets.p, ar.recs.data(), ar.d_nets.bytes, hipMemcpyHostToDevice));
    TRY(A(ar.d_wave_net, (uint64_t)ar.words * 4));
    TRY(A(ar.d_x, n * 8));
    TRY(A(ar.d_s, n * 4));
    TRY(A(ar.d_ring, (uint64_t)ar.D * n * 4));
    TRY(A(ar.d_prow, n));
    TRY(A(ar.d_ptab, 256 * sizeof(nh::PRow)));
    if (c.n_inputs) {
        TRY(A(ar.d_stim_ptr, n * 4));
        TRY(A(ar.d_stim_hi, n * 4));
        TRY(A(ar.d_stim_input, c.cap_stim * 4));
        TRY(A(ar.d_stim_w, c.cap_stim * 4));
        // column group of slot q: its request's inputs
        ar.stim_group.resize(c.max_requests);
        for (uint32_t q = 0; q < c.max_requests; ++q) ar.stim_group[q] = make_uint2(q * c.n_inputs, c.n_inputs);
        TRY(ar.alloc_stim(1024));
    } else {
        ar.stim_cap = 1024;
    }
    TRY(A(ar.d_spk, 4));
    TRY(A(ar.d_cnt, 16));
    TRY(A(ar.d_pn, n * 4));
    TRY(A(ar.d_last, n * 4));
    HIPCHK(hipMemset(ar.d_last.p, 0xFF, ar.d_last.bytes));
    TRY(A(ar.d_net_spikes, (uint64_t)c.max_requests * 8));
    TRY(A(ar.d_src_ptr, n * 4));
    TRY(A(ar.d_row_hi, n * 4));
    TRY(A(ar.d_syn, c.cap_synapses * 4));
    TRY(A(ar.d_w, c.cap_synapses * 4));
    TRY(A(ar.d_totals, 64));
    TRY(A(ar.d_rast_ctl, 32));
    if (c.n_windows) TRY(A(ar.d_wcnt, (uint64_t)c.n_windows * n * 4));
    if (s->pairs) {
        // per-wave table ranges start empty (free waves pull nothing)
        TRY(A(ar.d_pl_ptr, (uint64_t)ar.words * 4));
        TRY(A(ar.d_pl_hi, (uint64_t)ar.words * 4));
        TRY(A(ar.d_pl_ent, c.cap_synapses * 4));
        TRY(A(ar.d_pl_w, c.cap_synapses * 4));
        TRY(A(ar.d_pub, (uint64_t)ar.words * 16));
        TRY(A(ar.d_pair_abort, 4));
        HIPCHK(ar.h_pair_abort.alloc());
        ar.pair_tables = true;
        ar.fused = 2;
    }
    HIPCHK(hipDeviceSynchronize());
    ar.use_graph = 1;
    ar.graph_steps_req = c.chunk_steps;
    ar.set_stim_packer(pack_server, s);
    s->fn.init(n / 64);
    s->fs.init(c.cap_synapses);
    s->fst.init(c.cap_stim);
    s->fpl.init(s->pairs ? c.cap_synapses : 0);
    s->slot_req.assign(c.max_requests, nullptr);
    s->st.cap_neurons = n;
    s->st.cap_synapses = c.cap_synapses;
    update_grid(s);
    TRY(ar.run(0));  // fixes the round length (a multiple of D) and captures the graph
    // first-use costs off the first request's latency: the pinned staging
    // (sized for requests up to the capacities, at most 64 MiB; it grows if
    // a larger request comes) and the first launch of the range-clear kernel
    const uint64_t stage = std::min<uint64_t>(64ull << 20, c.cap_synapses * 8 + ar.n_total * 25 + (1u << 20));
    HIPCHK(s->stage.reserve(stage));
    nh::launch_clear_range(ar.args(), 0, 64, ar.stream);
    // and the first pinned host-to-device copies, large and small (into the
    // unused synapse area and the inactive slot records)
    const uint64_t big = std::min<uint64_t>(1u << 20, c.cap_synapses);
    uint32_t *z = s->stage.take<uint32_t>(big);
    std::memset(z, 0, big * 4);
    TRY(put(ar.d_syn, 0, z, big, ar.stream));
    nh::NetRec *r0 = s->stage.take<nh::NetRec>(1);
    *r0 = ar.recs[0];
    TRY(put(ar.d_nets, 0, r0, 1, ar.stream));
    HIPCHK(hipStreamSynchronize(ar.stream));
    return NH_OK;
}

// Device side of admission: upload the request into its ranges.
int admit(nh_server *s, Request *r) {
    Arena &ar = s->ar;
    hipStream_t st = ar.stream;
    // everything below is packed into the pinned stage (256-byte aligned
    // pieces) and copied asynchronously; the stream is synchronised at the end
    const uint64_t na = r->n_al;
    // (each piece rounds up to 256 B: 16 pieces of slack cover the fixed
    // ones, every new parameter row is a piece of its own)
    const size_t need = 256 * 20 + r->rows.size() * 256 + na * (1 + 8 + 8 + 8) + (size_t)r->E * 8 +
                        (size_t)r->n_stim * 8 + na / 64 * (2 + 8) + (size_t)r->E1 * 8 + sizeof(nh::NetRec);
    HIPCHK(s->stage.reserve(need));
    Stage &sg = s->stage;
    // parameter rows into the shared table
    std::vector<uint8_t> map_rows(r->rows.size());
    for (size_t q = 0; q < r->rows.size(); ++q) {
        size_t k = 0;
        for (; k < s->table.size(); ++k)
            if (std::memcmp(&s->table[k], &r->rows[q], sizeof(nh::PRow)) == 0) break;
        if (k == s->table.size()) {
            if (k >= 256) return set_error(NH_ERR_CAPACITY, "more than 256 distinct parameter rows");
            s->table.push_back(r->rows[q]);
            nh::PRow *pr = sg.take<nh::PRow>(1);
            *pr = r->rows[q];
            TRY(put(ar.d_ptab, k, pr, 1, st));
        }
        map_rows[q] = (uint8_t)k;
    }
    const uint64_t b = r->base;
    uint8_t *prow = sg.take<uint8_t>(na);
    std::memset(prow, 0, na);
    for (uint32_t i = 0; i < r->n; ++i) prow[i] = map_rows[r->row_of[i]];
    TRY(put(ar.d_prow, b, prow, na, st));
    uint32_t *lo = sg.take<uint32_t>(na), *hi = sg.take<uint32_t>(na);
    for (uint32_t i = 0; i < r->n_al; ++i) {
        const uint32_t a = i < r->n ? r->row_off[i] : r->row_off[r->n];
        const uint32_t z = i < r->n ? r->row_off[i + 1] : r->row_off[r->n];
        lo[i] = (uint32_t)(r->sbase + a);
        hi[i] = (uint32_t)(r->sbase + z);
    }
    TRY(put(ar.d_src_ptr, b, lo, na, st));
    TRY(put(ar.d_row_hi, b, hi, na, st));
    uint32_t *syn = sg.take<uint32_t>(r->E);
    for (uint64_t j = 0; j < r->E; ++j) {
        const uint32_t tg = (r->syn[j] & 0x07FFFFFFu) + (uint32_t)b;
        syn[j] = tg | (r->syn[j] & 0xF8000000u);
    }
    TRY(put(ar.d_syn, r->sbase, syn, r->E, st));
    uint32_t *wst = sg.take<uint32_t>(r->E);
    if (r->E) std::memcpy(wst, ar.ring_int ? (const void *)r->wq.data() : (const void *)r->wf.data(), r->E * 4);
    TRY(put(ar.d_w, r->sbase, wst, r->E, st));
    if (s->cfg.n_inputs) {
        uint32_t *slo = sg.take<uint32_t>(na), *shi = sg.take<uint32_t>(na);
        for (uint32_t i = 0; i < r->n_al; ++i) {
            const uint32_t a = i < r->n ? r->stim_off[i] : r->stim_off[r->n];
            const uint32_t z = i < r->n ? r->stim_off[i + 1] : r->stim_off[r->n];
            slo[i] = (uint32_t)(r->stbase + a);
            shi[i] = (uint32_t)(r->stbase + z);
        }
        TRY(put(ar.d_stim_ptr, b, slo, na, st));
        TRY(put(ar.d_stim_hi, b, shi, na, st));
        uint32_t *sin = sg.take<uint32_t>(r->n_stim);
        float *sw = sg.take<float>(r->n_stim);
        if (r->n_stim) {
            std::memcpy(sin, r->stim_in.data(), r->n_stim * 4);
            std::memcpy(sw, r->stim_w.data(), r->n_stim * 4);
        }
        TRY(put(ar.d_stim_input, r->stbase, sin, r->n_stim, st));
        TRY(put(ar.d_stim_w, r->stbase, sw, r->n_stim, st));
    }
    float2 *x0 = sg.take<float2>(na);
    std::fill(x0, x0 + na, make_float2(0.f, 0.f));
    std::copy(r->x0.begin(), r->x0.end(), x0);
    TRY(put(ar.d_x, b, x0, na, st));
    uint32_t *wn = sg.take<uint32_t>(na / 64);
    std::fill(wn, wn + na / 64, (uint32_t)r->slot);
    TRY(put(ar.d_wave_net, b / 64, wn, na / 64, st));
    if (s->pairs) {
        // delay-1 in-table: per-wave ranges in the pool, sources to arena ids
        // (base + local < 2^26, so no carry into the lane field)
        const uint64_t nw = na / 64;
        uint32_t *plo = sg.take<uint32_t>(nw), *phi = sg.take<uint32_t>(nw);
        uint64_t at = r->plbase;
        for (uint64_t w = 0; w < nw; ++w) {
            plo[w] = (uint32_t)at;
            at += r->pl_cnt[w];
            phi[w] = (uint32_t)at;
        }
        uint32_t *pe = sg.take<uint32_t>(r->E1), *pw = sg.take<uint32_t>(r->E1);
        for (uint64_t e = 0; e < r->E1; ++e) pe[e] = r->pl_ent[e] + (uint32_t)b;
        if (r->E1) std::memcpy(pw, r->pl_w.data(), r->E1 * 4);
        TRY(put(ar.d_pl_ptr, b / 64, plo, nw, st));
        TRY(put(ar.d_pl_hi, b / 64, phi, nw, st));
        TRY(put(ar.d_pl_ent, r->plbase, pe, r->E1, st));
        TRY(put(ar.d_pl_w, r->plbase, pw, r->E1, st));
    }
    // ring columns, s, counters of the range start at zero
    nh::StepArgs a = ar.args();
    nh::launch_clear_range(a, (uint32_t)b, r->n_al, st);
    // per-request record; its local step 0 is the next arena step
    nh::NetRec &rec = ar.recs[r->slot];
    std::memset(&rec, 0, sizeof(rec));
    rec.base = (uint32_t)b;
    rec.n = r->n;
    rec.seed = r->cfg.seed;
    rec.t0 = (int64_t)ar.t;
    Arena::fill_noise(rec, r->cfg);
    rec.int_scale = r->cfg.int_scale;
    rec.stim_base = r->slot * s->cfg.n_inputs;
    rec.active = 1;
    rec.dt = r->cfg.dt_ms;
    rec.win_steps = r->win_steps;
    // the request stops exactly after its own steps, whatever the round length
    rec.t_len = (uint32_t)r->total_steps;
    nh::NetRec *rs = sg.take<nh::NetRec>(1);
    *rs = rec;
    TRY(put(ar.d_nets, r->slot, rs, 1, st));
    HIPCHK(hipStreamSynchronize(st));
    r->t0 = ar.t;
    r->admit_ns = now_ns();
    return NH_OK;
}

// Launch only the occupied prefix of the arena (first fit keeps it low);
// the grid moves in 64k-neuron steps so the step graph is rarely recaptured.
void update_grid(nh_server *s) {
    uint64_t hwm = 0;
    for (Request *q : s->slot_req)
        if (q) hwm = std::max<uint64_t>(hwm, q->base + q->n_al);
    const uint64_t g = std::min<uint64_t>(std::max<uint64_t>((hwm + 65535) & ~uint64_t(65535), 65536),
                                          s->ar.n_total);
    if (g != s->ar.grid_limit) {
        s->ar.grid_limit = (uint32_t)g;
        s->ar.graph_dirty = true;
    }
}

bool try_admit(nh_server *s, Request *r) {
    uint32_t slot = 0;
    for (; slot < s->cfg.max_requests; ++slot)
        if (!s->slot_req[slot]) break;
    if (slot == s->cfg.max_requests) return false;
    uint64_t nb, sb = 0, stb = 0;
    if (!s->fn.alloc(r->n_al / 64, nb)) return false;
    if (!s->fs.alloc(r->E, sb)) {
        s->fn.release(nb, r->n_al / 64);
        return false;
    }
    if (!s->fst.alloc(r->n_stim, stb)) {
        s->fn.release(nb, r->n_al / 64);
        s->fs.release(sb, r->E);
        return false;
    }
    uint64_t plb = 0;
    if (s->pairs && !s->fpl.alloc(r->E1, plb)) {
        s->fn.release(nb, r->n_al / 64);
        s->fs.release(sb, r->E);
        s->fst.release(stb, r->n_stim);
        return false;
    }
    r->plbase = plb;
    r->slot = slot;
    r->base = nb * 64;
    r->sbase = sb;
    r->stbase = stb;
    const int rc = admit(s, r);
    if (rc) {
        r->error = rc;
        r->error_msg = "admission failed";
        r->done = true;
        r->done_ns = now_ns();
        s->fn.release(nb, r->n_al / 64);
        s->fs.release(sb, r->E);
        s->fst.release(stb, r->n_stim);
        if (s->pairs) s->fpl.release(plb, r->E1);
        s->cv_done.notify_all();
        return true;  // consumed (failed)
    }
    s->slot_req[slot] = r;
    s->active++;
    update_grid(s);
    s->st.admitted++;
    s->st.peak_active = std::max<uint64_t>(s->st.peak_active, s->active);
    return true;
}

int retire(nh_server *s, Request *r) {
    Arena &ar = s->ar;
    const uint32_t nw = s->cfg.n_windows;
    r->counts.assign((size_t)nw * r->n, 0);
    if (nw && r->n)
        HIPCHK(hipMemcpy2DAsync(r->counts.data(), (size_t)r->n * 4, ar.d_wcnt.as<uint32_t>() + r->base,
                                (size_t)ar.n_total * 4, (size_t)r->n * 4, nw, hipMemcpyDeviceToHost,
                                ar.stream));
    // spikes and events of the request from its per-neuron lifetime counts
    std::vector<uint32_t> pn(r->n);
    if (r->n)
        HIPCHK(hipMemcpyAsync(pn.data(), ar.d_pn.as<uint32_t>() + r->base, (size_t)r->n * 4,
                              hipMemcpyDeviceToHost, ar.stream));
    nh::NetRec &rec = ar.recs[r->slot];
    rec.active = 0;
    TRY(put(ar.d_nets, r->slot, &rec, 1, ar.stream));
    std::vector<uint32_t> empty;
    if (s->pairs) {
        // the range's waves pull nothing from now on
        empty.assign(r->n_al / 64, 0u);
        TRY(put(ar.d_pl_ptr, r->base / 64, empty.data(), empty.size(), ar.stream));
        TRY(put(ar.d_pl_hi, r->base / 64, empty.data(), empty.size(), ar.stream));
    }
    HIPCHK(hipStreamSynchronize(ar.stream));
    uint64_t sp = 0, ev = 0;
    for (uint32_t i = 0; i < r->n; ++i) {
        sp += pn[i];
        ev += (uint64_t)pn[i] * (r->row_off[i + 1] - r->row_off[i]);
    }
    r->spikes = sp;
    s->st.events += ev;
    s->fn.release(r->base / 64, r->n_al / 64);
    s->fs.release(r->sbase, r->E);
    s->fst.release(r->stbase, r->n_stim);
    if (s->pairs) s->fpl.release(r->plbase, r->E1);
    s->slot_req[r->slot] = nullptr;
    s->active--;
    update_grid(s);
    s->st.completed++;
    s->st.spikes += sp;
    r->done_ns = now_ns();
    r->done = true;
    return NH_OK;
}

void worker_loop(nh_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        while (!s->queue.empty()) {
            Request *r = s->queue.front();
            if (!try_admit(s, r)) break;
            s->queue.pop_front();
        }
        if (s->active == 0) {
            if (s->stop) break;
            s->cv_submit.wait(lk);
            continue;
        }
        lk.unlock();
        // one scheduling round: every in-flight request advances together
        const uint64_t k = s->ar.graph_steps ? s->ar.graph_steps : s->cfg.chunk_steps;
        int rc = s->ar.run(k);
        if (!rc) rc = s->ar.sync_timed();
        lk.lock();
        if (rc) {
            s->worker_error = rc;
            s->worker_msg = nh_last_error();
            for (auto *r : s->slot_req)
                if (r) {
                    r->error = rc;
                    r->done = true;
                }
            s->cv_done.notify_all();
            break;
        }
        s->st.steps += k;
        s->st.gpu_ms += s->ar.last_run_ms;
        if (s->ar.pair_usable()) s->st.paired_rounds++;
        for (uint32_t q = 0; q < s->cfg.max_requests; ++q) {
            Request *r = s->slot_req[q];
            if (r && s->ar.t - r->t0 >= r->total_steps) {
                if (retire(s, r)) {
                    r->error = NH_ERR_HIP;
                    r->done = true;
                }
            }
        }
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

void nh_server_config_default(nh_server_config *c) {
    std::memset(c, 0, sizeof(*c));
    c->max_requests = 256;
    c->max_delay = 31;
    c->n_inputs = 8;
    c->n_windows = 10;
    c->chunk_steps = 100;
    c->int_weights = 1;
    c->int_scale = 65536.0f;
    c->hbm_fraction = 0.25f;
}

int nh_server_create(const nh_server_config *cfg, nh_server **out) {
    return nh::guarded([&]() -> int {
        if (!cfg || !out) return set_error(NH_ERR_ARG, "null argument");
        *out = nullptr;
        int dev = 0;
        if (hipGetDeviceCount(&dev) != hipSuccess || dev < 1) return set_error(NH_ERR_NO_DEVICE, "no HIP device");
        if (cfg->max_delay < 1 || cfg->max_delay > 31) return set_error(NH_ERR_ARG, "max_delay must be 1..31");
        if (cfg->max_requests < 1 || cfg->max_requests > 65535) return set_error(NH_ERR_ARG, "max_requests");
        std::unique_ptr<nh_server> s(new (std::nothrow) nh_server());
        if (!s) return set_error(NH_ERR_CAPACITY, "out of host memory");
        s->cfg = *cfg;
        nh_server_config &c = s->cfg;
        if (c.chunk_steps == 0) c.chunk_steps = 100;
        // admission by free HBM: capacities left 0 come from the free memory
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const double budget = (double)fr * (c.hbm_fraction > 0.0f ? c.hbm_fraction : 0.25f);
        const double per_neuron = 8 + 4 + 4.0 * (c.max_delay + 1) + 1 + 8 + 8 + 8 + 4.0 * c.n_windows +
                                  (c.n_inputs ? 8 + 8 * 8 : 0);
        if (!c.cap_neurons) c.cap_neurons = std::min<uint64_t>((uint64_t)(budget * 0.5 / per_neuron), 16ull << 20);
        // per synapse: target word + weight, and the same again for the
        // delay-1 in-table pool of the paired shape
        if (!c.cap_synapses) c.cap_synapses = std::min<uint64_t>((uint64_t)(budget * 0.5 / 16.0), 1ull << 30);
        c.cap_neurons = std::min<uint64_t>(c.cap_neurons, 1ull << 27) & ~uint64_t(63);
        if (c.cap_neurons < 64) return set_error(NH_ERR_CAPACITY, "not enough HBM for the arena");
        if (!c.cap_stim) c.cap_stim = c.n_inputs ? c.cap_neurons : 0;
        TRY(build_arena(s.get()));
        s->worker = std::thread(worker_loop, s.get());
        *out = s.release();
        return NH_OK;
    });
}

void nh_server_destroy(nh_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        // requests still queued will not run
        s->queue.clear();
    }
    s->cv_submit.notify_all();
    if (s->worker.joinable()) s->worker.join();
    delete s;
}

int nh_server_submit(nh_server *s, const nh_net_desc *d, const nh_config *cfg, const float *stim,
                     uint64_t stim_steps, uint64_t total_steps, uint32_t win_steps, uint64_t *id) {
    return nh::guarded([&]() -> int {
        if (!s || !d || !cfg || !id) return set_error(NH_ERR_ARG, "null argument");
        TRY(nh_check_net(d, cfg));
        if ((cfg->int_weights ? 1 : 0) != s->cfg.int_weights)
            return set_error(NH_ERR_ARG, "request int_weights differs from the server's ring");
        if (d->n_inputs > s->cfg.n_inputs) return set_error(NH_ERR_ARG, "too many inputs for the server");
        if (d->n && d->n_state != 2) return set_error(NH_ERR_ARG, "state row width must be 2");
        if (d->n_params > (uint32_t)nh::kMaxParams) return set_error(NH_ERR_ARG, "param row too wide");
        if (total_steps == 0 || total_steps > 0xFFFFFFFFull)
            return set_error(NH_ERR_ARG, "total_steps must be 1..2^32-1");
        auto r = std::make_unique<Request>();
        r->submit_ns = now_ns();  // latency counts the request's preparation too
        r->n = d->n;
        r->n_al = (d->n + 63) & ~63u;
        r->cfg = *cfg;
        r->total_steps = total_steps;
        r->win_steps = win_steps;
        r->E = d->n ? d->src_ptr[d->n] : 0;
        if (r->n_al / 64 > s->cfg.cap_neurons / 64 || r->E > s->cfg.cap_synapses)
            return set_error(NH_ERR_CAPACITY, "request larger than the server arena");
        // rows sorted by (delay, target), local targets packed with the delay
        // (src_ptr is checked monotone from 0 by nh_check_net); large requests
        // split their rows over a few threads: the preparation is part of the
        // request's latency
        if (r->n) r->row_off.assign(d->src_ptr, d->src_ptr + r->n + 1);
        else r->row_off.assign(1, 0);
        r->syn.resize(r->E);
        if (s->cfg.int_weights) r->wq.resize(r->E);
        else r->wf.resize(r->E);
        {
            const bool iw = s->cfg.int_weights != 0;
            // chunks of about 32k synapses (at most 8): small requests stay on
            // the calling thread
            const uint32_t nt = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(8, r->E / 32768 + 1), r->n ? r->n : 1);
            std::vector<const char *> err(nt, nullptr);
            std::atomic<uint32_t> next{0};
            auto work = [&] {
                for (uint32_t q; (q = next.fetch_add(1)) < nt;)
                    err[q] = prep_rows(d, cfg, s->cfg.max_delay, iw, (uint32_t)((uint64_t)r->n * q / nt),
                                       (uint32_t)((uint64_t)r->n * (q + 1) / nt), r.get());
            };
            nh::run_workers(work, nt);
            for (const char *e : err)
                if (e) return set_error(std::strstr(e, "custom-tau") ? NH_ERR_UNSUPPORTED : NH_ERR_ARG, e);
        }
        if (s->pairs) {
            // delay-1 in-table: rows are sorted by delay, so a row's delay-1
            // synapses are its prefix; entries of a wave in source order
            const uint32_t nw = r->n_al / 64;
            r->pl_cnt.assign(nw, 0);
            for (uint32_t i = 0; i < r->n; ++i)
                for (uint32_t o = r->row_off[i]; o < r->row_off[i + 1] && (r->syn[o] >> 27) == 1u; ++o)
                    r->pl_cnt[(r->syn[o] & 0x07FFFFFFu) / 64]++;
            std::vector<uint32_t> at(nw + 1, 0);
            for (uint32_t w = 0; w < nw; ++w) at[w + 1] = at[w] + r->pl_cnt[w];
            r->E1 = at[nw];
            r->pl_ent.resize(r->E1);
            r->pl_w.resize(r->E1);
            for (uint32_t i = 0; i < r->n; ++i)
                for (uint32_t o = r->row_off[i]; o < r->row_off[i + 1] && (r->syn[o] >> 27) == 1u; ++o) {
                    const uint32_t tg = r->syn[o] & 0x07FFFFFFu, e = at[tg / 64]++;
                    r->pl_ent[e] = i | ((tg & 63u) << 26);
                    std::memcpy(&r->pl_w[e],
                                s->cfg.int_weights ? (const void *)&r->wq[o] : (const void *)&r->wf[o], 4);
                }
        }
        // distinct parameter rows
        r->row_of.resize(r->n);
        for (uint32_t i = 0; i < r->n; ++i) {
            if (d->model[i] >= d->n_models || d->model_ids[d->model[i]] > 1)
                return set_error(NH_ERR_ARG, "bad model index");
            nh::PRow pr{};
            pr.model = d->model_ids[d->model[i]];
            for (uint32_t q = 0; q < d->n_params; ++q) pr.p[q] = d->p[(uint64_t)i * d->n_params + q];
            size_t k = 0;
            for (; k < r->rows.size(); ++k)
                if (std::memcmp(&r->rows[k], &pr, sizeof(pr)) == 0) break;
            if (k == r->rows.size()) {
                if (k >= 256) return set_error(NH_ERR_CAPACITY, "too many distinct parameter rows");
                r->rows.push_back(pr);
            }
            r->row_of[i] = (uint8_t)k;
        }
        r->x0.resize(r->n);
        for (uint32_t i = 0; i < r->n; ++i) r->x0[i] = make_float2(d->x0[2 * (uint64_t)i], d->x0[2 * (uint64_t)i + 1]);
        r->stim_off.assign(r->n + 1, 0);
        if (d->n_inputs && d->stim_ptr) {
            for (uint32_t i = 0; i <= r->n; ++i) r->stim_off[i] = d->stim_ptr[i];
            r->n_stim = d->stim_ptr[r->n];
            r->stim_in.assign(d->stim_input, d->stim_input + r->n_stim);
            r->stim_w.assign(d->stim_w, d->stim_w + r->n_stim);
        }
        if (s->cfg.n_inputs) {
            r->stim_steps = stim_steps;
            r->stim.assign(stim_steps * s->cfg.n_inputs, 0.0f);
            for (uint64_t t = 0; t < stim_steps; ++t)
                for (uint32_t k = 0; k < d->n_inputs; ++k)
                    r->stim[t * s->cfg.n_inputs + k] = stim[t * d->n_inputs + k];
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->worker_error) return set_error(s->worker_error, "server stopped: " + s->worker_msg);
        r->id = s->next_id++;
        *id = r->id;
        s->queue.push_back(r.get());
        s->all.emplace(r->id, std::move(r));
        s->cv_submit.notify_all();
        return NH_OK;
    });
}

int nh_server_wait(nh_server *s, uint64_t id, uint32_t timeout_ms, uint32_t *counts, uint64_t cap,
                   nh_request_stats *st) {
    if (!s) return set_error(NH_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return set_error(NH_ERR_ARG, "unknown request id");
    Request *r = it->second.get();
    if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return r->done; }))
        return set_error(NH_ERR_NONE, "request not finished");
    if (r->error) return set_error(r->error, "request failed: " + r->error_msg);
    if (counts) {
        if (cap < r->counts.size()) return set_error(NH_ERR_CAPACITY, "counts buffer too small");
        std::copy(r->counts.begin(), r->counts.end(), counts);
    }
    if (st) {
        st->submit_ns = r->submit_ns;
        st->admit_ns = r->admit_ns;
        st->done_ns = r->done_ns;
        st->t0 = r->t0;
        st->spikes = r->spikes;
        st->n = r->n;
    }
    s->all.erase(it);
    return NH_OK;
}

int nh_server_get_stats(nh_server *s, nh_server_stats *out) {
    if (!s || !out) return set_error(NH_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;  // events and spikes of completed requests
    return NH_OK;
}

}  // extern "C"
