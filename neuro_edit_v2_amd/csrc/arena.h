// arena.h -- host side of the step engine: device buffers of one arena
// (one network, or many packed by offset tables), the per-chunk launch
// sequence (direct launches or a captured hipGraph), and readbacks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/neuro_hip.h"
#include "host_util.h"
#include "layout.h"

namespace nh {

int check_desc(const nh_net_desc &d, const nh_config &c, uint32_t *max_delay);
void launch_step(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool push,
                 bool packed, bool lean, hipStream_t st);
void launch_step_binned(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool packed,
                        hipStream_t st);
void launch_scatter(const StepArgs &a, uint32_t blocks, int ring, bool packed, hipStream_t st);
void launch_watch(const StepArgs &a, hipStream_t st);
void launch_advance(uint64_t *t_dev, uint32_t k, hipStream_t st);
void launch_copy(const void *src, void *dst, uint64_t bytes, uint32_t lane_bytes, hipStream_t st);
void launch_empty(uint32_t blocks, hipStream_t st);
void launch_pair(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool packed,
                 bool lean, bool w8, hipStream_t st);
uint32_t pair_resident_waves(int ring, int pmode, int mmode, bool packed, bool lean, bool w8,
                             uint32_t block);
void launch_run(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool packed, hipStream_t st);
uint32_t run_resident_waves(int ring, int pmode, int mmode, bool packed, uint32_t block);
void launch_count_rows(const uint64_t *bits, uint32_t R, uint32_t words, uint32_t base, uint32_t n,
                       uint64_t t, uint32_t rows, uint32_t *out, hipStream_t st);
// Words of the file-scope count scratch (nh_spike_counts_noalloc), and its
// device address on device dev (looked up once per device).
extern const uint32_t kScratchWords;
hipError_t scratch_on_device(int dev, uint32_t **ptr);

int set_error(int code, const std::string &msg);

// Owned device allocation.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        release();
        bytes = n ? n : 4;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            bytes = 0;
            return e;
        }
        e = hipMemset(p, 0, bytes);
        if (e != hipSuccess) return e;
        // the zero fill runs on the null stream, which the arenas'
        // non-blocking streams do not wait for: it is finished before the
        // buffer is handed out (a buffer allocated after creation is first
        // touched by work on those streams)
        return hipStreamSynchronize(nullptr);
    }
    template <class T> T *as() const { return (T *)p; }
};

// A 32-bit word of pinned host memory mapped into the device's address
// space: device code writes it rarely (a wait that gave up), and the host
// reads it after a synchronisation without a copy.
struct PinnedWord {
    uint32_t *host = nullptr;
    uint32_t *dev = nullptr;
    PinnedWord() = default;
    PinnedWord(const PinnedWord &) = delete;
    PinnedWord &operator=(const PinnedWord &) = delete;
    ~PinnedWord() {
        if (host) (void)hipHostFree(host);
    }
    hipError_t alloc() {
        if (host) return hipSuccess;
        hipError_t e = hipHostMalloc((void **)&host, 4, hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) {
            host = nullptr;
            return e;
        }
        *(volatile uint32_t *)host = 0u;
        e = hipHostGetDevicePointer((void **)&dev, host, 0);
        return e;
    }
    uint32_t get() const { return host ? *(volatile uint32_t *)host : 0u; }
    void clear() {
        if (host) *(volatile uint32_t *)host = 0u;
    }
};

// One packed network as the arena sees it.
struct NetSlot {
    uint32_t base = 0, n = 0;
    uint32_t n_inputs = 0, stim_base = 0;
    nh_config cfg{};
};

struct Tuning {
    int fused = -1;          // -1 auto, 0 two-phase, 1 fused, 2 paired (two steps per launch),
                             // 3 persistent (many steps per launch)
    int use_graph = -1;      // -1 auto
    uint32_t graph_steps = 64;
    uint32_t scatter_blocks = 0;  // 0 auto
    uint32_t block = 0;           // 0 = kBlock
    uint32_t lds_bins = 0;        // LDS-binned push (one-step shape): bins per workgroup, 0 = off
};

class Arena {
  public:
    ~Arena();
    // Build from n networks.  All networks share int_weights; the ring
    // depth is the max of theirs (a deeper ring leaves results unchanged).
    int build(uint32_t n_nets, const nh_net_desc *nets, const nh_config *cfgs, uint32_t raster_steps,
              const Tuning &tune);
    int reset();
    // Run n steps asynchronously (stimulus windows packed by the packer).
    int run(uint64_t n_steps);
    int sync();
    // sync(), then the last run's event times into last_run_ms and
    // last_tail_ms (read only when asked for: two event queries would
    // otherwise lengthen every synchronisation)
    int sync_timed();
    int upload_recs();
    static void fill_noise(NetRec &r, const nh_config &c);
    uint32_t neuron_blocks() const;
    uint32_t scat_blocks() const;
    int upload_stim(uint64_t t_arena, uint32_t k);
    // Stimulus window buffers for stim_cap steps of stim_cols columns, plus
    // the per-step "inputs not all zero" flag of each column group (a
    // network's inputs: group g = columns [stim_group[g].x, +stim_group[g].y)).
    int alloc_stim(uint32_t cap);
    std::vector<uint2> stim_group;
    DBuf d_stim_nz;

    // Stimulus: the caller supplies a packer producing [k][stim_cols].
    using StimPacker = void (*)(void *ctx, uint64_t t_arena, uint32_t k, float *out);
    void set_stim_packer(StimPacker p, void *ctx) { packer_ = p; packer_ctx_ = ctx; }

    // readbacks (synchronous)
    int read_x(std::vector<float2> &out);
    int read_f32(const DBuf &b, std::vector<float> &out, size_t n);
    int read_u32(const DBuf &b, std::vector<uint32_t> &out, size_t n);
    int read_u64(const DBuf &b, std::vector<uint64_t> &out, size_t n);
    int spikes_total(uint64_t &out);
    int spikes_in(uint32_t base, uint32_t n, uint64_t &out);
    int events_total(uint64_t &out);
    // Counter snapshot taken on the stream (a device copy of pn_total, no
    // host wait), and the spikes and events since it (synchronises).
    int mark_counters();
    int counts_since_mark(uint64_t &spikes, uint64_t &events);
    int drain_raster();   // device raster window -> host vectors
    // Per-neuron spike counts of arena neurons [base, base + n) over the
    // `rows` steps before t (0 < rows <= bits_rows, rows < t), counted on the
    // device into dev_out (n u32 of device memory) and copied to host_out;
    // synchronises.
    int count_rows(uint32_t base, uint32_t n, uint64_t rows, uint32_t *dev_out, uint32_t *host_out);
    // Lifetime spike counts of arena neurons [base, base + n) into host_out.
    int read_pn(uint32_t base, uint32_t n, uint32_t *host_out);
    DBuf d_counts;             // count_rows output (allocated at first use)

    // state
    uint32_t n_nets = 0, n_total = 0, D = 2, words = 0, bits_rows = 0;
    int ring_int = 0, pmode = 0, mmode = 0, packed = 1, fused = 0, use_graph = 0;
    uint64_t t = 0;            // arena steps done
    // The step counter d_t as it stands once the enqueued work has run.  The
    // step kernels read their step as d_t + k; direct launches without
    // stimulus inputs give it as an offset from this value (no counter
    // update per run), captured chunks and the rest bring d_t to their
    // first step and advance it by their length.
    uint64_t t_dev_val = 0;
    bool capturing = false;    // launch_chunk records into a graph
    uint64_t n_syn = 0;
    std::vector<NetSlot> slots;
    std::vector<NetRec> recs;
    hipStream_t stream = nullptr;
    int device = 0;            // the device current at creation
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_run_ms = 0.0;
    // Run time after its first step launch: ev_tail is recorded right behind
    // the run's first direct step launch (or its first graph replay), and
    // the step launches after it are counted, so that ev_tail -> ev1 over
    // that count is the back-to-back time per launch, without the host's
    // latency before the first kernel.
    hipEvent_t ev_tail = nullptr;
    bool tail_marking = false, tail_marked = false;
    uint64_t tail_launches = 0;
    double last_tail_ms = 0.0;
    uint64_t last_tail_launches = 0;
    int note_launches(uint64_t n);
    int begin_timing();
    int end_timing();
    uint32_t stim_cols = 0, stim_cap = 0;
    uint32_t n_watch = 0, xrec_depth = 4096;
    bool has_ext = false;
    uint64_t dev_bytes = 0;

    // raster (host side accumulation)
    uint32_t rast_steps = 0;                    // rows of the device raster window
    uint64_t rast_step0 = 0;                    // origin of the device window
    std::vector<uint32_t> h_rast_cnt;           // per step counts (drained)
    std::vector<uint32_t> h_rast_ids;           // ids (ascending per step)
    uint64_t h_rast_step0 = 0;                  // first step held on the host

    // device buffers
    DBuf d_t, d_nets, d_wave_net, d_x, d_x0, d_s, d_ring, d_ext, d_prow, d_ptab, d_p, d_mreg;
    DBuf d_stim_ptr, d_stim_input, d_stim_w, d_stim_win;
    DBuf d_spk, d_cnt, d_bits, d_pn, d_last, d_net_spikes;
    DBuf d_pn_mark;            // pn_total at the last mark_counters
    bool pn_marked = false;
    DBuf d_rast_bits, d_rast_ctl;
    DBuf d_src_ptr, d_syn, d_delay16, d_w, d_totals;
    DBuf d_watch, d_xrec;
    DBuf d_row_hi, d_wcnt;     // served arena: explicit row ends, window counters
    DBuf d_stim_hi;            // served arena: explicit stimulus range ends
    // custom-tau synapses
    DBuf d_xt_ptr, d_xt_idx, d_xdecay, d_xsyn, d_xring, d_xs_lo, d_xs_hi, d_xs_k, d_xs_d, d_xs_w;
    DBuf d_xargs;              // one XArgs naming the buffers above
    XArgs xargs_h{};           // its host copy
    uint32_t n_x = 0;
    // paired-step shape: delay-1 in-tables per wave, published spike words,
    // wait-abort flag; built when the shape may be used (the ring then has
    // one slot more than the deepest delay needs)
    DBuf d_pl_ptr, d_pl_hi, d_pl_ent, d_pl_w, d_pub, d_pair_abort;  // d_pl_hi: served arena
    PinnedWord h_pair_abort;   // host mirror of d_pair_abort
    bool pair_tables = false;
    bool pair_launched = false;          // paired launches since the last abort check
    bool persist_launched = false;       // of them, persistent launches
    uint64_t pair_fallbacks = 0;         // synchronisations that found a paired wait given up
    uint32_t pair_cap_block = 0, pair_cap_waves = 0, pair_cap8_waves = 0;
    bool pair_w8 = false;                // the next paired launches use the 8-wave build
    // test hook (failure handling): poll limit of the pair wait (0 = the
    // default), and whether wave 0 withholds its published word
    uint32_t pair_spins = 0, pair_mute = 0;
    // true when the next chunk may use the paired kernel (and which build:
    // pair_w8)
    bool pair_usable();
    // one network, no stimulus inputs, no custom-tau synapses, no window
    // counters: the step kernels' lean variants (fewer arguments read)
    bool lean() const { return n_nets == 1 && !dynamic && !d_stim_ptr.p && n_x == 0 && !d_wcnt.p; }
    // the one-step kernel's lean variant pays off only on large grids, where
    // the scalar-load throughput of many waves binds (measured: 0.17 us
    // slower per step at 1.6k waves, 0.17 us faster at 4.7k, 0.54 us at 15.6k)
    bool lean_step() const { return lean() && (uint64_t)neuron_blocks() * (block / 64) >= 3072; }
    int check_pair_abort();
    // persistent kernel (many steps per launch): published words kPubDepth
    // steps deep, per-residue group counters and done tags; usable when
    // fused == 3, the arena is lean, has the delay-1 tables and its grid is
    // co-resident
    DBuf d_pub3, d_gctr, d_done;
    uint32_t run_groups = 1;
    uint32_t run_cap_block = 0, run_cap_waves = 0;
    uint32_t run_max_steps = 2048;      // steps per persistent launch at most
    bool run_usable();
    std::vector<uint32_t> h_deg_;      // total out-degree per arena neuron
    uint32_t n_win = 0;
    bool dynamic = false;      // served arena: slots come and go
    uint32_t grid_limit = 0;   // served arena: launch only the occupied prefix
    uint32_t block = kBlock;   // workgroup size of the step kernels
    uint32_t lds_bins = 0;     // LDS-binned push: bins per workgroup (0 = wave-by-wave push)
    bool binned() const { return fused == 1 && lds_bins != 0; }
    // Time each (shape, workgroup, graph chunk) candidate over trial_steps
    // from the initial state, keep the fastest, then reset (t must be 0).
    int autotune(uint32_t trial_steps, std::string &report);
    float *h_stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    int stage_i = 0;
    bool stage_used[2] = {false, false};

    StepArgs args() const;
    // The device copy of args() that kernels taking a pointer read (k and
    // slot zeroed): refreshed at the start of every run (and profile) when
    // it differs, after the stream has drained (a kernel in flight may read
    // the old copy).
    DBuf d_args;
    StepArgs args_dev_{};
    bool args_dev_valid = false;
    int sync_dev_args();
    int launch_chunk(uint64_t t0, uint32_t k);
    int launch_persistent(uint64_t t0, uint32_t k);
    int check_args(const StepArgs &a) const;
    // direct launches with an event pair around every step kernel
    int profile(uint64_t n_steps, double &main_ms, uint64_t &main_n, double &aux_ms,
                uint64_t &aux_n);
    int capture_graph();
    int sync_t_dev(uint64_t to);
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;

    StimPacker packer_ = nullptr;
    void *packer_ctx_ = nullptr;
    uint32_t P_ = 4;
    bool graph_dirty = true;
    uint32_t graph_steps = 0, graph_steps_req = 64;
    uint32_t scatter_blocks = 0;
    std::vector<PRow> krows_;
    bool timing_valid = false;
    std::vector<uint32_t> h_src_ptr_;   // host copy of the arena CSR row starts

  private:
    int alloc(DBuf &b, size_t bytes);
    template <class T> int upload(DBuf &b, const std::vector<T> &v);
};

}  // namespace nh
