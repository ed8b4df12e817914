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
// Device-memory pointers of the kernel arguments.  In device code they carry
// the global address space, so loads, stores and atomics through pointers
// read out of memory (the argument block's device copy) compile to global
// memory operations instead of flat ones (which count against the scalar
// and LDS counters too, so every wait for a scalar load would also wait for
// them).  Host code (and the host sources' device passes, which hold no
// kernels) sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__) && defined(NH_DEVICE_KERNELS)
template <class T> using gptr = __attribute__((address_space(1))) T *;
#else
template <class T> using gptr = T *;
#endif

struct XArgs {
    uint32_t n_x;
    gptr<const uint32_t> xt_ptr;      // n_total+1
    gptr<const uint32_t> xt_idx;
    gptr<const float> xdecay;
    gptr<float> xsyn;
    gptr<void> xring;                 // D x n_x (float or int32)
    gptr<const uint32_t> xs_lo;
    gptr<const uint32_t> xs_hi;
    gptr<const uint32_t> xs_k;
    gptr<const uint16_t> xs_d;
    gptr<const void> xs_w;
};

// Kernel arguments of one step.  Pointers may be null when a feature is off.
// Steps of one launch chunk at most (the step kernels take a launch's step
// offset within its chunk in 12 bits).
constexpr uint32_t kMaxChunkSteps = 4096;

struct StepArgs {
    uint32_t n_total;            // arena neurons (multiple of 64)
    uint32_t block;              // workgroup size (64..512, multiple of 64)
    uint32_t D;                  // ring depth (max over packed networks)
    gptr<const uint64_t> t_dev;       // arena step of chunk start (device)
    const StepArgs *self;        // the arena's device copy of these arguments
                                 // (all fields but k and slot, which the
                                 // kernels that read the copy take apart)
    uint32_t k;                  // step offset within the chunk
    uint32_t slot;               // (t_dev + k) % D, fixed by the launcher
    uint32_t single_net;         // 1: every wave uses net0 (no table read)
    gptr<const uint32_t> wave_net;    // per wave: index into nets (a word: read with a scalar load)
    gptr<const NetRec> nets;
    NetRec net0;                 // the network of a single-network arena
    // state
    gptr<float2> x;                   // (v, u|w) per neuron
    gptr<float> s;                    // decaying synaptic current
    gptr<void> ring;                  // D x n_total (float or int32)
    uint32_t ring_bytes;         // ring size when it is below 2 GB (buffer atomics)
    uint32_t ring_buf;           // 1: deliveries use buffer atomics over ring_bytes
    gptr<const float> ext;            // external current or null
    // parameters
    uint32_t n_krows;            // rows in krow (1..kKargRows) in kernel-argument mode
    PRow krow[kKargRows];        // parameter rows carried in the arguments
    gptr<const uint8_t> prow;         // per-neuron row index (table modes)
    gptr<const PRow> ptab;            // global table (<= 256 rows)
    gptr<const float> p;              // per-neuron rows (n_total x P) in row mode
    gptr<const uint8_t> mreg;         // per-neuron registry model id (row mode)
    uint32_t P;
    // stimulus
    gptr<const uint32_t> stim_lo;     // per-neuron stimulus entry range or null
    gptr<const uint32_t> stim_hi;
    gptr<const uint32_t> stim_input;
    gptr<const float> stim_w;
    gptr<const float> stim_win;       // [chunk step][columns]
    gptr<const uint32_t> stim_nz;     // [chunk step][group]: 0 = every input of the group is zero
    uint32_t stim_groups;
    uint32_t stim_cols;
    // spike outputs
    gptr<uint32_t> spk;               // per-step spike lists (3 x n_total)
    gptr<uint32_t> cnt;               // per-list counters
    gptr<uint64_t> bits;              // record ring (rows x words) or null
    uint32_t bits_rows;
    uint32_t words;
    gptr<uint32_t> pn_total;          // per-neuron lifetime spikes
    gptr<uint32_t> last_spike;        // per-neuron last spike step (kNoSpike = never)
    gptr<uint64_t> net_spikes;        // per-network spike totals or null
    // raster stream
    gptr<uint64_t> rast_bits;         // raster window: [step - origin][words] spike words, or null
    gptr<uint64_t> rast_ctl;          // [1] step origin of the window, [2] overflow flag
    uint32_t rast_cap_steps;     // rows of rast_bits
    // windowed spike counters [window][n_total] (served form) or null
    gptr<uint32_t> wcnt;
    uint32_t n_win;
    // synapses (CSR by source, rows sorted by delay then target); the row
    // of neuron i is [row_lo[i], row_hi[i]) -- for a packed CSR row_hi is
    // row_lo + 1 (the same array), the served arena keeps both
    gptr<const uint32_t> row_lo;
    gptr<const uint32_t> row_hi;
    gptr<const uint32_t> syn;         // packed (tgt | delay << 27) or tgt
    gptr<const uint16_t> delay16;     // unpacked delays (null when packed)
    gptr<const void> w;               // float or int32 per synapse
    gptr<uint64_t> totals;            // spare counters (cleared with the arena)
    gptr<const XArgs> xt;             // custom-tau synapses, or null (none)
    // paired-step kernel (two steps per launch): per wave, the delay-1
    // synapses into the wave ([pl_ptr[w], pl_hi[w]) of pl_ent = source |
    // target lane << 26, weight pl_w; pl_hi = pl_ptr + 1 for a packed
    // table, the served arena keeps both); the spike word each wave
    // publishes for the first step (pub[2w + half] = tag << 32 | half
    // word); the flag a wave raises when its bounded wait runs out
    gptr<const uint32_t> pl_ptr;
    gptr<const uint32_t> pl_hi;
    gptr<const uint32_t> pl_ent;
    gptr<const void> pl_w;
    gptr<uint64_t> pub;
    gptr<uint32_t> pair_abort;
    gptr<uint32_t> pair_abort_host;   // the same flag in mapped host memory (read by the host without a copy)
    uint32_t pair_spins;         // polls before a wave gives up waiting
    uint32_t pair_mute;          // test only: 1 = wave 0 starts late (paired) or
                                 // never publishes (persistent)
    // persistent kernel (many steps per launch): steps of the launch; the
    // spike words it publishes, kPubDepth steps deep (pub3[(t % 3)][2w +
    // half], tag t + 1); per step residue r = t % 4 and group g of
    // workgroups (workgroup b is in group b % n_groups), the group's arrival
    // counter gctr[r][g] and its done tag done[r][g] = t + 1, published once
    // every workgroup of the group has finished step t (pushes complete)
    uint32_t run_steps;
    gptr<uint64_t> pub3;
    gptr<uint32_t> gctr;
    gptr<uint32_t> done;
    uint32_t n_groups;           // 1..64
    // LDS-binned push (k_step_binned): entries of the workgroup's bin table
    // (a power of two; 0 = the wave-by-wave push)
    uint32_t bin_entries;
    // watched neurons (record ring of state rows)
    gptr<const uint32_t> watch;
    uint32_t n_watch;
    gptr<float2> xrec;
    uint32_t xrec_depth;
};

}  // namespace nh
