// layout.h -- device-memory layout shared by the host code and the kernels.
//
// One "arena" of neurons holds one network (plain path) or many networks
// packed back to back (batched / served path).  Every network's neuron range
// starts on a 64-neuron boundary, so a wave64 always covers neurons of one
// network and its spike ballot is exactly one 64-bit word of the spike
// bitmap.  Per-network scalars live in a NetRec table indexed per wave.
#pragma once
#include <stdint.h>

namespace nh {

constexpr int kWave = 64;
constexpr int kBlock = 256;          // default workgroup: 4 waves
constexpr int kMaxBlock = 512;       // largest workgroup the step kernels accept
constexpr int kMaxParams = 10;       // AdEx row width
constexpr uint32_t kNoSpike = 0xFFFFFFFFu;
constexpr uint32_t kPairSpinsDefault = 1u << 16;  // polls before a paired-step wave gives up
constexpr uint32_t kPubDepth = 3;    // persistent kernel: steps of published spike words kept

// Per-network scalars (one row per packed network).  All fields are read
// with scalar loads (the index is wave-uniform).
struct alignas(64) NetRec {
    uint32_t base;        // first arena neuron of the network
    uint32_t n;           // neurons of the network
    uint64_t seed;        // noise seed
    int64_t t0;           // arena step at which the network's local step is 0
    float pkick;          // rate*dt/1000 (Poisson-kick probability)
    float kick;           // kick_mv/dt (0 when kick_mv == 0)
    float sd;             // Gaussian current sd (0 = off)
    float decay_s;        // exp(-dt/tau_syn)
    float int_scale;      // int ring scale
    uint32_t stim_base;   // first column of this network in the stimulus window
    uint32_t active;      // 0 = slot empty (served form)
    float dt;             // step, ms
    uint32_t win_steps;   // per-window spike counters: steps per window (0 = off)
    uint32_t t_len;       // local steps the network runs (0 = unbounded)
};
static_assert(sizeof(NetRec) == 64, "NetRec is one cache line");

// Parameter table row (type-indexed parameters: most networks have a few
// distinct rows, e.g. RS and FS).
struct alignas(16) PRow {
    float p[kMaxParams];
    uint32_t model;       // registry id: 0 izh, 1 adex
    uint32_t pad;
};
static_assert(sizeof(PRow) == 48, "PRow size");

constexpr int kKargRows = 4;        // parameter rows carried in the kernel arguments

// Custom-tau synapses (own exponential trace, sim/src/state.rs:230-274): per
// slot k a trace xsyn[k] with decay xdecay[k] and an arrival ring
// xring[slot][k]; target-major lists xt for the fold-in, per-source rows xs
// (k, delay, w) for the push.  Kept in device memory behind one pointer of
// the step arguments: every wave reads its arguments with scalar loads, and
// a group behind a pointer that may be null is read only where it is used.
struct XArgs {
    uint32_t n_x;
    const uint32_t *xt_ptr;      // n_total+1
    const uint32_t *xt_idx;
    const float *xdecay;
    float *xsyn;
    void *xring;                 // D x n_x (float or int32)
    const uint32_t *xs_lo;
    const uint32_t *xs_hi;
    const uint32_t *xs_k;
    const uint16_t *xs_d;
    const void *xs_w;
};

// Kernel arguments of one step.  Pointers may be null when a feature is off.
// Steps of one launch chunk at most (the step kernels take a launch's step
// offset within its chunk in 12 bits).
constexpr uint32_t kMaxChunkSteps = 4096;

struct StepArgs {
    uint32_t n_total;            // arena neurons (multiple of 64)
    uint32_t block;              // workgroup size (64..512, multiple of 64)
    uint32_t D;                  // ring depth (max over packed networks)
    const uint64_t *t_dev;       // arena step of chunk start (device)
    const StepArgs *self;        // the arena's device copy of these arguments
                                 // (all fields but k and slot, which the
                                 // kernels that read the copy take apart)
    uint32_t k;                  // step offset within the chunk
    uint32_t slot;               // (t_dev + k) % D, fixed by the launcher
    uint32_t single_net;         // 1: every wave uses net0 (no table read)
    const uint16_t *wave_net;    // per wave: index into nets
    const NetRec *nets;
    NetRec net0;                 // the network of a single-network arena
    // state
    float2 *x;                   // (v, u|w) per neuron
    float *s;                    // decaying synaptic current
    void *ring;                  // D x n_total (float or int32)
    uint32_t ring_bytes;         // ring size when it is below 2 GB (buffer atomics)
    uint32_t ring_buf;           // 1: deliveries use buffer atomics over ring_bytes
    const float *ext;            // external current or null
    // parameters
    uint32_t n_krows;            // rows in krow (1..kKargRows) in kernel-argument mode
    PRow krow[kKargRows];        // parameter rows carried in the arguments
    const uint8_t *prow;         // per-neuron row index (table modes)
    const PRow *ptab;            // global table (<= 256 rows)
    const float *p;              // per-neuron rows (n_total x P) in row mode
    const uint8_t *mreg;         // per-neuron registry model id (row mode)
    uint32_t P;
    // stimulus
    const uint32_t *stim_lo;     // per-neuron stimulus entry range or null
    const uint32_t *stim_hi;
    const uint32_t *stim_input;
    const float *stim_w;
    const float *stim_win;       // [chunk step][columns]
    const uint32_t *stim_nz;     // [chunk step][group]: 0 = every input of the group is zero
    uint32_t stim_groups;
    uint32_t stim_cols;
    // spike outputs
    uint32_t *spk;               // per-step spike lists (3 x n_total)
    uint32_t *cnt;               // per-list counters
    uint64_t *bits;              // record ring (rows x words) or null
    uint32_t bits_rows;
    uint32_t words;
    uint32_t *pn_total;          // per-neuron lifetime spikes
    uint32_t *last_spike;        // per-neuron last spike step (kNoSpike = never)
    uint64_t *net_spikes;        // per-network spike totals or null
    // raster stream
    uint64_t *rast_bits;         // raster window: [step - origin][words] spike words, or null
    uint64_t *rast_ctl;          // [1] step origin of the window, [2] overflow flag
    uint32_t rast_cap_steps;     // rows of rast_bits
    // windowed spike counters [window][n_total] (served form) or null
    uint32_t *wcnt;
    uint32_t n_win;
    // synapses (CSR by source, rows sorted by delay then target); the row
    // of neuron i is [row_lo[i], row_hi[i]) -- for a packed CSR row_hi is
    // row_lo + 1 (the same array), the served arena keeps both
    const uint32_t *row_lo;
    const uint32_t *row_hi;
    const uint32_t *syn;         // packed (tgt | delay << 27) or tgt
    const uint16_t *delay16;     // unpacked delays (null when packed)
    const void *w;               // float or int32 per synapse
    uint64_t *totals;            // spare counters (cleared with the arena)
    const XArgs *xt;             // custom-tau synapses, or null (none)
    // paired-step kernel (two steps per launch): per wave, the delay-1
    // synapses into the wave ([pl_ptr[w], pl_hi[w]) of pl_ent = source |
    // target lane << 26, weight pl_w; pl_hi = pl_ptr + 1 for a packed
    // table, the served arena keeps both); the spike word each wave
    // publishes for the first step (pub[2w + half] = tag << 32 | half
    // word); the flag a wave raises when its bounded wait runs out
    const uint32_t *pl_ptr;
    const uint32_t *pl_hi;
    const uint32_t *pl_ent;
    const void *pl_w;
    uint64_t *pub;
    uint32_t *pair_abort;
    uint32_t *pair_abort_host;   // the same flag in mapped host memory (read by the host without a copy)
    uint32_t pair_spins;         // polls before a wave gives up waiting
    uint32_t pair_mute;          // test only: 1 = wave 0 never publishes
    // persistent kernel (many steps per launch): steps of the launch; the
    // spike words it publishes, kPubDepth steps deep (pub3[(t % 3)][2w +
    // half], tag t + 1); per step residue r = t % 4 and group g of
    // workgroups (workgroup b is in group b % n_groups), the group's arrival
    // counter gctr[r][g] and its done tag done[r][g] = t + 1, published once
    // every workgroup of the group has finished step t (pushes complete)
    uint32_t run_steps;
    uint64_t *pub3;
    uint32_t *gctr;
    uint32_t *done;
    uint32_t n_groups;           // 1..64
    // LDS-binned push (k_step_binned): entries of the workgroup's bin table
    // (a power of two; 0 = the wave-by-wave push)
    uint32_t bin_entries;
    // watched neurons (record ring of state rows)
    const uint32_t *watch;
    uint32_t n_watch;
    float2 *xrec;
    uint32_t xrec_depth;
};

}  // namespace nh
