// kernels.hip -- the per-step hot path for gfx950 (wave64, CDNA4).
//
// Step t of the reference CPU backend (sim/src/cpu.rs:192-306) is: read and
// clear ring slot t, integrate every neuron, then scatter the spikes of
// step t into slots t+delay (delay >= 1).  The scatter of step t never
// touches slot t, and the update of step t reads no slot but t, so one
// kernel per step suffices:
//
//  * k_step<PUSH=true> ("fused" shape): each wave updates its 64 neurons,
//    ballots its spikes (the wave64 mask is one word of the spike bitmap),
//    then pushes the out-synapse rows of its own spiking neurons into the
//    ring with hardware atomics: the rows are walked as one concatenated
//    list, every load of a 256-entry batch issued before its atomics.  No
//    cross-wave dependency exists inside a step, so there is no grid
//    barrier; the kernel boundary orders step t's pushes before step t+1's
//    reads.
//
//  * k_step<PUSH=false> + k_scatter ("two-kernel" shape): the update
//    compacts the spikes of the step into a list; a second kernel scatters
//    one spike per wave.  Kept as the reference shape and for comparison.
//
//  * k_pair ("paired" shape): two steps per launch where the whole grid is
//    resident at once; the delay-1 arrivals of the first step are pulled by
//    their targets' waves from published spike words (see its comment).
//    k_pair8 is the same code built for 8 waves per SIMD, for one network
//    whose grid is too large for k_pair's 6.
//
// The int ring (sim/src/kernels.rs:57-66) sums integers, which commute, so
// every shape gives the CPU raster bit for bit whatever the atomic order.
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#define NH_DEVICE_KERNELS 1  // device pointers of the arguments in the global address space
#include "device.h"
#include "layout.h"

namespace nh {

// Raw buffer atomics (LLVM intrinsics): an access whose offset is outside
// the descriptor's range is dropped by the hardware, which lets a batch of
// deliveries skip its empty slots without a branch per slot.
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
__device__ int32_t buf_atomic_add_i32(int32_t v, i32x4 rsrc, int32_t voff, int32_t soff,
                                      int32_t aux) __asm("llvm.amdgcn.raw.buffer.atomic.add.i32");
__device__ float buf_atomic_fadd_f32(float v, i32x4 rsrc, int32_t voff, int32_t soff,
                                     int32_t aux) __asm("llvm.amdgcn.raw.buffer.atomic.fadd.f32");
constexpr uint32_t kBufDrop = 0x80000000u;  // out-of-range offset: the access is dropped

// Raw (unstrided) buffer descriptor over `bytes` bytes at p (gfx9 family:
// 32-bit data format in the third dword).
__device__ __forceinline__ i32x4 buffer_rsrc(const void *p, uint32_t bytes) {
    const uint64_t a = (uint64_t)p;
    i32x4 r;
    r.x = (int32_t)(uint32_t)a;
    r.y = (int32_t)(uint32_t)(a >> 32);
    r.z = (int32_t)bytes;
    r.w = 0x00020000;
    return r;
}

template <int RING> struct Ring;
// Ring read scale: x / scale (sim/src/kernels.rs:60), or x * 2^-k when the
// scale is 2^k with 0 <= k <= 64 -- both exact then, so the same bits.
struct RingScale {
    float scale, inv;
    bool pow2;
    __device__ __forceinline__ explicit RingScale(float s) : scale(s) {
        const uint32_t b = __float_as_uint(s);
        pow2 = (b & 0x7FFFFFu) == 0u && b >= 0x3F800000u && b <= 0x5F800000u;
        inv = __uint_as_float(0x7F000000u - b);
    }
};

template <> struct Ring<0> {
    using T = float;
    static __device__ __forceinline__ float get(T v, const RingScale &) { return v; }
    static __device__ __forceinline__ void add(T *p, T v) { unsafeAtomicAdd(p, v); }
    static __device__ __forceinline__ void add_buf(i32x4 r, uint32_t off, T v) {
        (void)buf_atomic_fadd_f32(v, r, (int32_t)off, 0, 0);
    }
};
template <> struct Ring<1> {
    using T = int32_t;
    // sim/src/kernels.rs:60: (float)x / WSCALE
    static __device__ __forceinline__ float get(T v, const RingScale &rs) {
        if (rs.pow2) return (float)v * rs.inv;
        // a real (uniform) branch: keeps the division off the common path
        asm volatile("");
        return (float)v / rs.scale;
    }
    static __device__ __forceinline__ void add(T *p, T v) { atomicAdd(p, v); }
    static __device__ __forceinline__ void add_buf(i32x4 r, uint32_t off, T v) {
        (void)buf_atomic_add_i32(v, r, (int32_t)off, 0, 0);
    }
};

__device__ __forceinline__ uint32_t wave_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// A wave-uniform value that no kernel writes while a step kernel runs (step
// counter, per-wave table bounds, network records), read through the
// constant address space so that it is a scalar load whatever precedes it
// (the compiler proves a plain load unclobbered only up to the first
// instruction that may write memory): a wave's vector loads complete in
// order, so a vector load of it issued behind the lanes' state loads could
// be used only once they had all returned.
// Wait here for the load of v (an empty asm that reads it).  Where a value
// loaded long before is first read inside a loop, the wait the compiler
// places there also covers whatever the loop's other paths leave
// outstanding (in the paired kernel: the spike word just published, loads
// just issued); read once here, the loop needs no wait for it.
__device__ __forceinline__ void ready(uint32_t v) { asm volatile("" ::"v"(v)); }

// A step output written through to memory (WT): an agent-scope relaxed
// atomic store of the same value (a global store with the write-through
// bit).  The end of a launch then has few dirty L2 lines left to write back.
// Measured: the paired kernel at 100k neurons 5.36 -> 5.00 us per launch,
// the 256-network batch 26.4 -> 25.5 us per step; the lean one-step kernel
// at 1M neurons, which streams, 9.0 -> 9.3 us, so it keeps plain stores
// (WT false).  Nontemporal stores, which stay in L2, changed nothing.
template <bool WT = true, typename T>
__device__ __forceinline__ void store_wt(T *p, T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "word or double word");
    if constexpr (!WT) {
        *p = v;
    } else if constexpr (sizeof(T) == 4) {
        uint32_t b;
        __builtin_memcpy(&b, &v, 4);
        __hip_atomic_store((uint32_t *)p, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        uint64_t b;
        __builtin_memcpy(&b, &v, 8);
        __hip_atomic_store((uint64_t *)p, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <typename T>
__device__ __forceinline__ T sload(gptr<const T> p) {
    static_assert(sizeof(T) % 4 == 0, "scalar loads read whole words");
    if constexpr (sizeof(T) <= 8) {
        return *(const __attribute__((address_space(4))) T *)p;
    } else {
        // a record: word by word (one scalar load of each run of words)
        const __attribute__((address_space(4))) uint32_t *q = (const __attribute__((address_space(4))) uint32_t *)p;
        uint32_t w[sizeof(T) / 4];
#pragma unroll
        for (uint32_t j = 0; j < sizeof(T) / 4; ++j) w[j] = q[j];
        T r;
        __builtin_memcpy(&r, w, sizeof(T));
        return r;
    }
}

// Entries pushed per lane before the first atomic of a batch: every load of
// a batch is in flight at once, so a batch costs one memory round trip.
// A wave that gives up a bounded wait raises the abort flag: in device
// memory for the other waves' checks, and in mapped host memory for the
// host's check after the launch.
__device__ __forceinline__ void raise_abort(const StepArgs &a) {
    __hip_atomic_store(a.pair_abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a.pair_abort_host) __hip_atomic_store(a.pair_abort_host, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// s_waitcnt immediate for vmcnt(0) with the other counters left alone
// (gfx9 encoding: vmcnt[3:0] and [15:14], expcnt[6:4], lgkmcnt[11:8]).
constexpr uint32_t kWaitVmcnt0 = 0x0F70;

constexpr uint32_t kPushUnroll = 4;
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;  // batch slot without an entry

// Ring slot and target of one delivery: synapse word sy (target and delay
// packed in one word, or target plus a separate 16-bit delay d16).
template <bool PACKED>
__device__ __forceinline__ uint64_t ring_index(const StepArgs &a, uint32_t sy, uint32_t d16,
                                               uint32_t slot0) {
    uint32_t tgt, d;
    if (PACKED) {
        tgt = sy & 0x07FFFFFFu;
        d = sy >> 27;
    } else {
        tgt = sy;
        d = d16;
    }
    uint32_t sl = slot0 + d;
    sl = sl >= a.D ? sl - a.D : sl;
    return (uint64_t)sl * a.n_total + tgt;
}

// One batch of a push: up to kPushUnroll entries per lane (kNoEntry = none),
// their synapse words and weights once loaded.
template <typename RT>
struct PushBatch {
    uint32_t j[kPushUnroll];
    uint32_t sy[kPushUnroll], d16[kPushUnroll];
    RT wv[kPushUnroll];
};

// Load a batch's entries (unconditional: an empty slot reads entry 0).
template <int RING, bool PACKED>
__device__ __forceinline__ void load_batch(const StepArgs &a, PushBatch<typename Ring<RING>::T> &pb) {
    using RT = typename Ring<RING>::T;
    const RT *w = (const RT *)a.w;
#pragma unroll
    for (uint32_t r = 0; r < kPushUnroll; ++r) {
        const uint32_t jr = pb.j[r] != kNoEntry ? pb.j[r] : 0u;
        pb.sy[r] = a.syn[jr];
        pb.wv[r] = w[jr];
        pb.d16[r] = PACKED ? 0u : a.delay16[jr];
    }
}

// Deliver a loaded batch.  With a ring below 2 GB the atomics are buffer
// atomics whose empty slots carry an out-of-range offset, so the batch has
// no branch per slot (branches there make the compiler wait for each atomic
// before the next); otherwise each slot's atomic is guarded.  SKIP1 drops
// the delay-1 entries (the paired-step kernel pulls those instead).
template <int RING, bool PACKED, bool SKIP1 = false>
__device__ __forceinline__ void send_batch(const StepArgs &a, const PushBatch<typename Ring<RING>::T> &pb,
                                           uint32_t slot0) {
    using RT = typename Ring<RING>::T;
    bool keep[kPushUnroll];
#pragma unroll
    for (uint32_t r = 0; r < kPushUnroll; ++r)
        keep[r] = pb.j[r] != kNoEntry && (!SKIP1 || (PACKED ? pb.sy[r] >> 27 : pb.d16[r]) != 1u);
    if (a.ring_buf) {
        const i32x4 rs = buffer_rsrc(a.ring, a.ring_bytes);
#pragma unroll
        for (uint32_t r = 0; r < kPushUnroll; ++r) {
            // ring below 2 GB: the byte offset fits in 31 bits
            const uint32_t off = (uint32_t)ring_index<PACKED>(a, pb.sy[r], pb.d16[r], slot0) * 4u;
            Ring<RING>::add_buf(rs, keep[r] ? off : kBufDrop, pb.wv[r]);
        }
    } else {
#pragma unroll
        for (uint32_t r = 0; r < kPushUnroll; ++r)
            if (keep[r])
                Ring<RING>::add((RT *)a.ring + ring_index<PACKED>(a, pb.sy[r], pb.d16[r], slot0), pb.wv[r]);
    }
}

// Load the entries j[0..kPushUnroll) and deliver them: all loads first,
// then the atomics.
template <int RING, bool PACKED, bool SKIP1 = false>
__device__ __forceinline__ void deliver_batch(const StepArgs &a, const uint32_t (&j)[kPushUnroll],
                                              uint32_t slot0) {
    PushBatch<typename Ring<RING>::T> pb;
#pragma unroll
    for (uint32_t r = 0; r < kPushUnroll; ++r) pb.j[r] = j[r];
    load_batch<RING, PACKED>(a, pb);
    send_batch<RING, PACKED, SKIP1>(a, pb, slot0);
}

// Push one spike's out-synapse row into the ring; lanes stride the row
// (delay-sorted, so a run of lanes hits one ring row with ascending targets),
// kPushUnroll strides per batch.
template <int RING, bool PACKED>
__device__ __forceinline__ void push_row(const StepArgs &a, uint32_t lo, uint32_t hi,
                                         uint32_t slot0, uint32_t lane) {
    for (uint32_t j0 = lo + lane; j0 < hi; j0 += kPushUnroll * kWave) {
        uint32_t j[kPushUnroll];
#pragma unroll
        for (uint32_t r = 0; r < kPushUnroll; ++r) {
            const uint32_t jr = j0 + r * kWave;
            j[r] = jr < hi ? jr : kNoEntry;
        }
        deliver_batch<RING, PACKED>(a, j, slot0);
    }
}

// A wave's push as one list: the rows of its spiking lanes concatenated in
// lane order, walked kPushUnroll*64 entries at a time, so a wave whose
// spikes own few synapses pays one load round trip in total instead of one
// per spike and per 64 entries.  `live`: spikes whose rows are not fully
// issued (always a suffix of the spikes in lane order); `live_start`: list
// offset of the first.
struct PushList {
    uint64_t live;
    uint32_t live_start, total;
};

__device__ __forceinline__ PushList push_list(uint64_t mask, uint32_t row_lo, uint32_t row_hi) {
    uint32_t total = 0;
    for (uint64_t m = mask; m;) {
        const uint32_t b = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        total += __builtin_amdgcn_readlane(row_hi, b) - __builtin_amdgcn_readlane(row_lo, b);
    }
    return PushList{mask, 0u, total};
}

// The entries of list positions [base, base + kPushUnroll*64) into j.
__device__ __forceinline__ void list_batch(PushList &pl, uint32_t base, uint32_t row_lo, uint32_t row_hi,
                                           uint32_t lane, uint32_t (&j)[kPushUnroll]) {
    const uint32_t end = base + kPushUnroll * kWave;
#pragma unroll
    for (uint32_t r = 0; r < kPushUnroll; ++r) j[r] = kNoEntry;
    uint32_t s0 = pl.live_start;
    for (uint64_t m = pl.live; m;) {
        if (s0 >= end) break;
        const uint32_t b = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const uint32_t lo = __builtin_amdgcn_readlane(row_lo, b);
        const uint32_t d = __builtin_amdgcn_readlane(row_hi, b) - lo;
#pragma unroll
        for (uint32_t r = 0; r < kPushUnroll; ++r) {
            const uint32_t u = base + r * kWave + lane - s0;
            if (u < d) j[r] = lo + u;
        }
        if (s0 + d <= end) {
            pl.live &= ~(1ull << b);
            pl.live_start = s0 + d;
        }
        s0 += d;
    }
}

// Push the list's batches from position `from` on.  PIPE: pipelined, the
// next batch's loads issued before the current batch's atomics (a wave's
// loads complete in order, so waiting for the current batch does not wait
// for the next one's).  Measured: the paired kernel in the reference's
// bench regime (≈ 500 deliveries per wave and step) 6.34 -> 5.91 us per
// step; the batched one-step kernel 1.5 % slower, so it keeps the plain loop.
template <int RING, bool PACKED, bool SKIP1 = false, bool PIPE = false>
__device__ __forceinline__ void push_list_from(const StepArgs &a, PushList &pl, uint32_t from,
                                               uint32_t row_lo, uint32_t row_hi, uint32_t slot0,
                                               uint32_t lane) {
    using RT = typename Ring<RING>::T;
    if constexpr (!PIPE) {
        for (uint32_t base = from; base < pl.total; base += kPushUnroll * kWave) {
            uint32_t j[kPushUnroll];
            list_batch(pl, base, row_lo, row_hi, lane, j);
            deliver_batch<RING, PACKED, SKIP1>(a, j, slot0);
        }
    } else {
        if (from >= pl.total) return;
        PushBatch<RT> cur;
        list_batch(pl, from, row_lo, row_hi, lane, cur.j);
        load_batch<RING, PACKED>(a, cur);
        for (uint32_t base = from + kPushUnroll * kWave;; base += kPushUnroll * kWave) {
            const bool more = base < pl.total;
            PushBatch<RT> nxt;
            if (more) {
                list_batch(pl, base, row_lo, row_hi, lane, nxt.j);
                load_batch<RING, PACKED>(a, nxt);
            }
            send_batch<RING, PACKED, SKIP1>(a, cur, slot0);
            if (!more) break;
            cur = nxt;
        }
    }
}

// The list of a wave's push and its first batch's entries; a wave with one
// spike (the common case) takes them without walking the mask.
__device__ __forceinline__ PushList push_first(uint64_t mask, uint32_t row_lo, uint32_t row_hi, uint32_t lane,
                                               uint32_t (&j)[kPushUnroll]) {
    if ((mask & (mask - 1u)) == 0ull) {
        const uint32_t b = (uint32_t)__builtin_ctzll(mask);
        const uint32_t lo = __builtin_amdgcn_readlane(row_lo, b), hi = __builtin_amdgcn_readlane(row_hi, b);
#pragma unroll
        for (uint32_t r = 0; r < kPushUnroll; ++r) {
            const uint32_t jr = lo + r * kWave + lane;
            j[r] = jr < hi ? jr : kNoEntry;
        }
        const uint32_t d = hi - lo;
        const bool done = d <= kPushUnroll * kWave;
        return PushList{done ? 0ull : mask, done ? d : 0u, d};
    }
    PushList pl = push_list(mask, row_lo, row_hi);
    list_batch(pl, 0u, row_lo, row_hi, lane, j);
    return pl;
}

// Push the rows of every spiking lane of a wave (PIPE: batches pipelined).
template <int RING, bool PACKED, bool SKIP1 = false, bool PIPE = false>
__device__ __forceinline__ void push_wave(const StepArgs &a, uint64_t mask, uint32_t row_lo,
                                          uint32_t row_hi, uint32_t slot0, uint32_t lane) {
    if constexpr (!PIPE) {
        uint32_t j[kPushUnroll];
        PushList pl = push_first(mask, row_lo, row_hi, lane, j);
        deliver_batch<RING, PACKED, SKIP1>(a, j, slot0);
        push_list_from<RING, PACKED, SKIP1>(a, pl, kPushUnroll * kWave, row_lo, row_hi, slot0, lane);
    } else {
        PushList pl = push_list(mask, row_lo, row_hi);
        push_list_from<RING, PACKED, SKIP1, true>(a, pl, 0u, row_lo, row_hi, slot0, lane);
    }
}

// Custom-tau synapses of one spike: arrival marks in the per-synapse ring
// (cpu.rs:271-273 pushes the synapse id; here the weight lands in its slot).
template <int RING>
__device__ __forceinline__ void push_custom(const StepArgs &a, uint32_t src, uint32_t slot0,
                                            uint32_t lane) {
    using RT = typename Ring<RING>::T;
    const XArgs &x = *a.xt;
    RT *xr = (RT *)x.xring;
    const RT *w = (const RT *)x.xs_w;
    const uint32_t lo = x.xs_lo[src], hi = x.xs_hi[src];
    for (uint32_t e = lo + lane; e < hi; e += kWave) {
        uint32_t sl = slot0 + x.xs_d[e];
        sl = sl >= a.D ? sl - a.D : sl;
        Ring<RING>::add(xr + (size_t)sl * x.n_x + x.xs_k[e], w[e]);
    }
}

// Noise of one neuron at local step tl: whether the Poisson kick fires and
// the Gaussian draw (0 when off).  Depends on the step, the neuron and the
// seed only, so a kernel may compute it before the step's input is known.
struct Noise {
    bool kick;
    float g;
};
__device__ __forceinline__ Noise noise_terms(const NetRec &nr, uint32_t local, uint64_t tl) {
    Noise z{false, 0.0f};
    if (nr.pkick > 0.0f || nr.sd != 0.0f) {
        // one hash serves the kick (high half) and the Gaussian (low half)
        const uint64_t key = noise_key(local, tl, nr.seed);
        const uint64_t h0 = mix64(key);
        z.kick = nr.pkick > 0.0f && kick_fires(h0, nr.pkick);
        if (nr.sd != 0.0f) z.g = gauss_draw(h0, key);
    }
    return z;
}

// Parameter row of neuron i (row index prow): kernel-argument rows
// (PMODE 0), a global row table (2) or per-neuron rows (1).
struct Params {
    float p[kMaxParams];
    uint32_t model;
};
// BFI: rows selected with one v_bfi_b32 per parameter (see below); otherwise
// with a select per parameter, which the compiler schedules more freely.
// Measured: the bit inserts are faster in the paired kernel (100k: 2.725
// against 2.745 us per step) and in the one-step kernel of the mixed
// Izhikevich/AdEx batch (24.62 against 25.02 us), slower in the
// Izhikevich-only one-step kernel (1M: 10.46 against 10.18 us; 100k one step
// per launch: 3.69 against 3.56 us).
template <int PMODE, int MMODE, bool BFI>
__device__ __forceinline__ Params select_params(const StepArgs &a, uint32_t i, uint32_t prow) {
    Params r_;
    r_.model = (uint32_t)MMODE;
    if (PMODE == 0) {
#pragma unroll
        for (int q = 0; q < kMaxParams; ++q) r_.p[q] = a.krow[0].p[q];
        r_.model = a.krow[0].model;
#pragma unroll
        for (int r = 1; r < kKargRows; ++r) {
            if ((uint32_t)r >= a.n_krows) break;  // uniform: only the rows in use
            if (!BFI) {
                const bool sel = prow == (uint32_t)r;
#pragma unroll
                for (int q = 0; q < kMaxParams; ++q) r_.p[q] = sel ? a.krow[r].p[q] : r_.p[q];
                r_.model = sel ? a.krow[r].model : r_.model;
                continue;
            }
            // One v_bfi_b32 per parameter, (m & row) | (~m & current), with
            // the row value read from its scalar register.  A select takes a
            // move and a v_cndmask per parameter (the lane mask already uses
            // the instruction's one scalar operand), and the compiler turns
            // the plain-C bit select back into that pair.
            const uint32_t m = prow == (uint32_t)r ? 0xFFFFFFFFu : 0u;
#pragma unroll
            for (int q = 0; q < kMaxParams; ++q) {
                uint32_t o;
                asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(o) : "v"(m), "s"(__float_as_uint(a.krow[r].p[q])),
                    "v"(__float_as_uint(r_.p[q])));
                r_.p[q] = __uint_as_float(o);
            }
            uint32_t om;
            asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(om) : "v"(m), "s"(a.krow[r].model), "v"(r_.model));
            r_.model = om;
        }
    } else if (PMODE == 2) {
        const PRow &pt = a.ptab[prow];
#pragma unroll
        for (int q = 0; q < kMaxParams; ++q) r_.p[q] = pt.p[q];
        r_.model = pt.model;
    } else {
        r_.model = MMODE == 2 ? a.mreg[i] : (uint32_t)MMODE;
        if (MMODE == 0 && a.P == 4) {
            const float4 v = *(const float4 *)(a.p + (size_t)i * 4);
            r_.p[0] = v.x; r_.p[1] = v.y; r_.p[2] = v.z; r_.p[3] = v.w;
        } else {
            for (uint32_t q = 0; q < a.P && q < (uint32_t)kMaxParams; ++q)
                r_.p[q] = a.p[(size_t)i * a.P + q];
        }
    }
    return r_;
}

// One neuron's update at step t (local step tl), in the order of
// sim/src/cpu.rs:217-260: ring arrival rv, custom-tau traces (ring slot
// `slot`), stimulus (window row kk, skipped when stim_on is false), Poisson
// kick and Gaussian current, model step.  With PRE the noise comes in nz
// (computed earlier by the caller); otherwise it is computed here, after
// the parameter rows.  s holds the synaptic current before the step and
// returns it decayed; xs is the model state.  Returns whether the neuron
// spiked.  LEAN: the arena has no custom-tau synapses (their pointers are
// never read).  HP: the parameter row comes in pp (selected once by the
// caller); otherwise it is selected here.
template <int RING, int PMODE, int MMODE, bool PRE, bool LEAN, bool HP>
__device__ __forceinline__ bool neuron_update(const StepArgs &a, const NetRec &nr, uint32_t i,
                                              uint32_t local, uint64_t tl, uint32_t slot,
                                              uint32_t kk, bool stim_on, uint32_t prow,
                                              typename Ring<RING>::T rv, float ext,
                                              const Noise &nz, const Params &pp, float &s,
                                              float2 &xs) {
    using RT = typename Ring<RING>::T;
    bool spiked;
    const RingScale rsc(nr.int_scale);
    const float arr = Ring<RING>::get(rv, rsc);
    const float sv = s + arr;
    float I = sv + ext;
    if (!LEAN && a.xt) {
        // custom-tau traces into this neuron (cpu.rs:208-228): this
        // step's arrival, use, then decay -- each slot has one target,
        // so its thread owns it
        const XArgs &x = *a.xt;
        RT *xr = (RT *)x.xring + (size_t)slot * x.n_x;
        for (uint32_t j = x.xt_ptr[i]; j < x.xt_ptr[i + 1]; ++j) {
            const uint32_t kx = x.xt_idx[j];
            const RT av = xr[kx];
            xr[kx] = (RT)0;
            const float xv = x.xsyn[kx] + Ring<RING>::get(av, rsc);
            I += xv;
            x.xsyn[kx] = xv * x.xdecay[kx];
        }
    }
    if (stim_on) {
        // Stimulus (stim.rs:119-127).  Skipped when every input of the
        // network is zero this step: each term is then w * 0 = +-0, and
        // I + (+-0) == I because I is never -0 here (ring sums start
        // from +0 and never reach -0, so s + arrival is not -0, nor is
        // that plus ext or plus a custom-tau trace).
        const float *row = a.stim_win + (size_t)kk * a.stim_cols + nr.stim_base;
        for (uint32_t j = a.stim_lo[i]; j < a.stim_hi[i]; ++j)
            I += a.stim_w[j] * row[a.stim_input[j]];
    }
    const Params pp_ = HP ? pp : select_params<PMODE, MMODE, MMODE != 0>(a, i, prow);
    const float *pr = pp_.p;
    const uint32_t model = pp_.model;
    const bool is_adex = MMODE == 1 || (MMODE == 2 && model == 1);
    // sim/src/cpu.rs:240-245: kick * model gain (C for AdEx, adex.rs:189-191)
    const float gain = is_adex ? pr[0] : 1.0f;
    if (PRE) {
        if (nz.kick) I += nr.kick * gain;
        if (nr.sd != 0.0f) I += (nr.sd * nz.g) * gain;
    } else if (nr.pkick > 0.0f || nr.sd != 0.0f) {
        // one hash serves the kick (high half) and the Gaussian (low half)
        const uint64_t key = noise_key(local, tl, nr.seed);
        const uint64_t h0 = mix64(key);
        if (nr.pkick > 0.0f && kick_fires(h0, nr.pkick)) I += nr.kick * gain;
        if (nr.sd != 0.0f) I += (nr.sd * gauss_draw(h0, key)) * gain;
    }
    if (is_adex)
        spiked = adex_step(pr, xs, I, nr.dt);
    else if (MMODE == 2 && model == 2)
        spiked = lif_step(pr, xs, I, nr.dt);   // noise gain 1 (the trait's default)
    else
        spiked = izh_step(pr[0], pr[1], pr[2], pr[3], xs, I, nr.dt);
    s = sv * nr.decay_s;
    return spiked;
}

// Per-neuron marks of a spike at step t (local step tl): lifetime count,
// last spike step, window counter.  No-return atomics: nothing waits on a
// read of the old count (each neuron owns its counter, so no contention).
// LEAN: no window counters.
template <bool LEAN>
__device__ __forceinline__ void spike_marks(const StepArgs &a, const NetRec &nr, uint32_t i,
                                            uint64_t t, uint64_t tl) {
    atomicAdd(&a.pn_total[i], 1u);
    a.last_spike[i] = t < 0xFFFFFFFEull ? (uint32_t)t : 0xFFFFFFFEu;
    if (!LEAN && a.wcnt && nr.win_steps) {
        const uint64_t wi = tl / nr.win_steps;
        if (wi < a.n_win) atomicAdd(&a.wcnt[(size_t)wi * a.n_total + i], 1u);
    }
}

// The wave's spike word of step t into the record ring and the raster
// window (lane 0 only).  The host drains the raster window before it
// fills; the flag reports a step outside it.
__device__ __forceinline__ void spike_words(const StepArgs &a, uint32_t wave, uint64_t t,
                                            uint64_t rast_origin, uint64_t mask) {
    if (a.bits) a.bits[(size_t)(t % a.bits_rows) * a.words + wave] = mask;
    if (a.rast_bits) {
        const uint64_t rs = t - rast_origin;
        if (rs < a.rast_cap_steps) a.rast_bits[(size_t)rs * a.words + wave] = mask;
        else a.rast_ctl[2] = 1;
    }
}

// ---------------------------------------------------------------- step
// PMODE: 0 parameter rows in the kernel arguments (<= 4 rows), 1 per-neuron
// rows, 2 global row table.  MMODE: 0 izh, 1 adex, 2 mixed.
// PUSH: the fused shape (own-spike push); otherwise the step appends its
// spikes to the per-step list that k_scatter reads.  LEAN: one network, no
// stimulus inputs, no custom-tau synapses, no window counters -- the
// pointers of those features are never read (every wave reads its
// arguments with scalar loads, whose throughput a launch of many waves
// feels).
// Packed arenas of several networks: a launch whose grid is a multiple of
// kBands workgroups visits the arena in kBands interleaved bands (workgroup
// b takes block (b % kBands) * (grid / kBands) + b / kBands), so that the
// workgroups in flight on a SIMD come from networks far apart in the arena
// -- a mix of models (an AdEx wave is VALU-heavy, an Izhikevich wave is
// memory-bound) instead of one network's waves side by side.  33 is coprime
// to the 8 XCDs the workgroups are dealt to, so every XCD sees every band
// (measured on the 256-network workload, us per step: 24.6 in arena order;
// 9 bands 24.2, 17 23.5, 33 23.3, 65 24.1).  Any grid size gives the same
// results: the map is a permutation of the grid's workgroups.
constexpr uint32_t kBands = 33;
constexpr bool kLateRows = true;

// Leading-argument packing of the step kernels: the wave count and the
// waves per workgroup in one word (wave count < 2^29, 1..8 waves), and the
// parameter-row index array only where a kernel variant reads it.
constexpr uint32_t kWpbShift = 29;
constexpr uint32_t kWordsMask = (1u << kWpbShift) - 1u;
// k_step's slot / step-offset / waves-per-workgroup word: slot < 2^17 (the
// ring holds at most 65,537 slots), step offset < 2^12 (kMaxChunkSteps)
constexpr uint32_t kSlotBits = 17;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;
constexpr uint32_t kKMask = (1u << (kWpbShift - kSlotBits)) - 1u;
static inline uint32_t pack_words_wpb(const StepArgs &a) { return a.words | (a.block / 64u - 1u) << kWpbShift; }
static inline const uint8_t *prow_arg(const StepArgs &a, int pmode) {
    return (pmode == 2 || (pmode == 0 && a.n_krows > 1)) ? a.prow : nullptr;
}

// The argument block of the step kernels: the lean builds (one network, one
// per-launch block that never changes) read the arena's device copy through
// a pointer -- a direct launch's own block is written by the host for every
// launch, so its first read misses every cache, while the copy stays in L2
// (measured: 100k paired launch 6.37 -> 5.94 us, 1M one-step 9.69 -> 9.13
// us).  The packed builds take the block by value: they run from captured
// graphs, whose blocks stay cached too, and read many of its fields (the
// copy measured 26.1 -> 27.6-28.8 us per batched step).
template <bool BY_POINTER>
using ArgBlock = typename std::conditional<BY_POINTER, const StepArgs *__restrict__, StepArgs>::type;
__device__ __forceinline__ const StepArgs &arg_block(const StepArgs *__restrict__ p) { return *p; }
__device__ __forceinline__ const StepArgs &arg_block(const StepArgs &a) { return a; }
template <bool BY_POINTER>
static inline ArgBlock<BY_POINTER> arg_pass(const StepArgs &a) {
    if constexpr (BY_POINTER) return a.self;
    else return a;
}

// The first loads' pointers and indices come as leading scalar arguments,
// preloaded into scalar registers (as for k_pair); band_len: workgroups per
// band when the packed arena's grid is visited in interleaved bands, else 0.
template <int RING, int PMODE, int MMODE, bool PUSH, bool PACKED, bool LEAN>
__global__ __launch_bounds__(kMaxBlock) void k_step(void *ring_p, float *s_p, float2 *x_p, uint32_t n_total,
                                                    uint32_t lk, uint32_t band_len, const uint8_t *prow_p,
                                                    ArgBlock<LEAN> av) {
    using RT = typename Ring<RING>::T;
    const StepArgs &a = arg_block(av);
    constexpr bool LATE_ROWS = LEAN && kLateRows;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t words = n_total >> 6;
    const uint32_t slot = lk & kSlotMask, k = (lk >> kSlotBits) & kKMask, wpb = (lk >> kWpbShift) + 1u;
    uint32_t blk = blockIdx.x;
    if (!LEAN && PUSH && band_len) blk = (blk % kBands) * band_len + blk / kBands;
    const uint32_t wave = blk * wpb + uniform(threadIdx.x >> 6);
    if (wave >= words) return;
    const uint32_t i = wave * kWave + lane;
    // ---- first batch of independent loads, issued before the network
    // record's (dependent) loads below
    RT *ring = (RT *)ring_p;
    const size_t ri = (size_t)slot * n_total + i;
    const RT rv = ring[ri];
    const float s_old = s_p[i];
    float2 xs = x_p[i];
    uint32_t prow = 0;
    if (PMODE != 1) {
        // parameter-row index, loaded without a branch (where there is no
        // index array a lane reads a byte of its own state and ignores it):
        // a load inside a branch is waited for where the branch ends, and
        // with it every load issued before it
        const uint8_t *pb = prow_p ? prow_p + i : (const uint8_t *)(x_p + i);
        const uint32_t r = *pb;
        prow = prow_p ? r : 0u;
    }
    // The network record is copied by value, so its fields are read once,
    // up front, into scalar registers (a reference that may point at either
    // the kernel arguments or the table makes every field a separate vector
    // load at its point of use).
    NetRec nr;
    uint32_t net = 0;
    if (LEAN || a.single_net) {
        nr = a.net0;
    } else {
        net = sload(a.wave_net + wave);
        nr = sload(a.nets + net);
    }
    const float ext = a.ext ? a.ext[i] : 0.0f;
    // CSR row bounds: loaded with the state, except in the lean variant (one
    // large network), where only a wave that spikes loads them, after its
    // ballot -- at 1M neurons about 3 % of the waves, 4 B per neuron-step
    // of HBM traffic saved against one more load round trip for those waves
    uint32_t row_lo = 0, row_hi = 0;
    if (PUSH && !LATE_ROWS) {
        row_lo = a.row_lo[i];
        row_hi = a.row_hi[i];
    }
    // (the step is read from the device counter even where the host knows
    // it: given by value, it measured slower -- the noise hashes then get
    // scheduled among the first loads)
    const uint64_t t = sload(a.t_dev) + k;
    // raster window origin (a scalar load issued with the others)
    const uint64_t rast_origin = a.rast_bits ? sload<uint64_t>(a.rast_ctl + 1) : 0;
    // does this network have a nonzero input this step?  (scalar load,
    // issued with the others)
    const bool stim_on = !LEAN && a.stim_lo && sload(a.stim_nz + (size_t)k * a.stim_groups + net) != 0u;

    const uint32_t local = i - nr.base;
    const uint64_t tl = t - (uint64_t)nr.t0;
    const bool act = local < nr.n && nr.active && (nr.t_len == 0 || tl < nr.t_len);
    bool spiked = false;
    float s_new = s_old;
    if (act)
        spiked = neuron_update<RING, PMODE, MMODE, false, LEAN, false>(
            a, nr, i, local, tl, slot, k, stim_on, prow, rv, ext, Noise{false, 0.0f}, Params{}, s_new, xs);

    // ---- wave64 ballot: the mask IS this wave's word of the spike bitmap
    const uint64_t mask = __ballot(spiked);
    if (PUSH && mask) {
        if (LATE_ROWS) {
            row_lo = a.row_lo[i];
            row_hi = a.row_hi[i];
        }
        // The push goes first, before any store of this step: a wave's
        // vector memory operations complete in order, so rows read after
        // the stores would wait for the stores.  The push writes slots
        // t+1..t+D-1 only, never the slot cleared below.
        push_wave<RING, PACKED>(a, mask, row_lo, row_hi, slot, lane);
        if (!LEAN && a.xt) {
            // custom-tau rows (rare): one spike at a time
            for (uint64_t m = mask; m;) {
                const uint32_t b = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                push_custom<RING>(a, wave * kWave + b, slot, lane);
            }
        }
    }
    // ---- this step's stores
    if (act) {
        store_wt<!LEAN>(&ring[ri], (RT)0);
        store_wt<!LEAN>(&x_p[i], xs);
        store_wt<!LEAN>(&s_p[i], s_new);
        if (spiked) spike_marks<LEAN>(a, nr, i, t, tl);
    }
    if (lane == 0) spike_words(a, wave, t, rast_origin, mask);
    const uint32_t L = (uint32_t)(t % 3);
    // two-kernel shape: the list of step t+1 starts empty (its appends come
    // in the next launch, after this one has finished)
    if (!PUSH && blockIdx.x == 0 && threadIdx.x == 0) a.cnt[(L + 1) % 3] = 0;
    if (!mask) return;
    // Spike and event totals are not counted here: they follow exactly from
    // pn_total and the out-degrees on readback, so no wave contends for a
    // shared counter.
    if (!PUSH) {
        const uint32_t nsp = (uint32_t)__popcll(mask);
        uint32_t off = 0;
        if (lane == 0) off = atomicAdd(&a.cnt[L], nsp);
        off = uniform(off);
        if (spiked) a.spk[(size_t)L * a.n_total + off + wave_rank(mask)] = i;
    }
}

// ---------------------------------------------------------------- binned
// The fused step with the push done by the whole workgroup through LDS
// bins (a packed arena, not the lean single-network case).  After the
// update each wave appends its spiking neurons' rows to a workgroup queue
// in LDS (row start and offset in the concatenated list, reserved with one
// 64-bit LDS atomic per wave, so offsets ascend with the queue index) and
// does its own stores; after a barrier a workgroup with spikes clears its
// table and every thread of the workgroup walks the concatenated list (entry -> owning row by binary search over
// the queue), loads the synapse, and adds its weight into an LDS hash table
// keyed by ring word (linear probing, at most kBinProbes slots; an entry
// that finds no slot goes straight to HBM); after a second barrier the
// table's words are added to the ring with one HBM atomic each.  Deliveries
// of the workgroup that land on the same ring word (same target, same
// arrival step) leave the workgroup as one atomic, and the row walk is
// spread over the workgroup's lanes instead of the spiking wave's.  Int
// ring: integer sums, so the raster stays bit-exact; float ring: the sum
// order changes (as with atomic order anyway).
constexpr uint32_t kBinEmpty = 0xFFFFFFFFu;
constexpr uint32_t kBinProbes = 8;

template <int RING, bool PACKED>
__device__ __forceinline__ void ring_add_word(const StepArgs &a, uint32_t word, typename Ring<RING>::T v) {
    using RT = typename Ring<RING>::T;
    if (a.ring_buf) Ring<RING>::add_buf(buffer_rsrc(a.ring, a.ring_bytes), word * 4u, v);
    else Ring<RING>::add((RT *)a.ring + word, v);
}

template <int RING, int PMODE, int MMODE, bool PACKED>
__global__ __launch_bounds__(kMaxBlock) void k_step_binned(StepArgs a) {
    using RT = typename Ring<RING>::T;
    extern __shared__ uint32_t smem[];
    const uint32_t B = a.bin_entries;
    uint32_t *bkey = smem;
    RT *bval = (RT *)(smem + B);
    uint32_t *q_lo = smem + 2u * B;               // blockDim entries
    uint32_t *q_off = q_lo + blockDim.x;          // blockDim entries
    unsigned long long *q_ctr = (unsigned long long *)(q_off + blockDim.x);  // spikes << 32 | entries
    if (threadIdx.x == 0) *q_ctr = 0ull;
    __syncthreads();

    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t blk = blockIdx.x;
    if (!a.single_net && gridDim.x % kBands == 0u) blk = (blk % kBands) * (gridDim.x / kBands) + blk / kBands;
    const uint32_t wave = blk * (blockDim.x >> 6) + uniform(threadIdx.x >> 6);
    const bool live = wave < a.words;              // every thread reaches the barriers
    const uint32_t i = wave * kWave + lane;
    const uint32_t slot = a.slot;
    uint64_t mask = 0;
    if (live) {
        NetRec nr;
        uint32_t net = 0;
        if (a.single_net) {
            nr = a.net0;
        } else {
            net = sload(a.wave_net + wave);
            nr = sload(a.nets + net);
        }
        RT *ring = (RT *)a.ring;
        const size_t ri = (size_t)slot * a.n_total + i;
        const RT rv = ring[ri];
        const float s_old = a.s[i];
        float2 xs = a.x[i];
        const float ext = a.ext ? a.ext[i] : 0.0f;
        const uint32_t row_lo = a.row_lo[i], row_hi = a.row_hi[i];
        uint32_t prow = 0;
        if (PMODE != 1 && (PMODE == 2 || a.n_krows > 1)) prow = a.prow[i];
        const uint64_t t = sload(a.t_dev) + a.k;
        const uint64_t rast_origin = a.rast_bits ? sload<uint64_t>(a.rast_ctl + 1) : 0;
        const bool stim_on = a.stim_lo && sload(a.stim_nz + (size_t)a.k * a.stim_groups + net) != 0u;
        const uint32_t local = i - nr.base;
        const uint64_t tl = t - (uint64_t)nr.t0;
        const bool act = local < nr.n && nr.active && (nr.t_len == 0 || tl < nr.t_len);
        bool spiked = false;
        float s_new = s_old;
        if (act)
            spiked = neuron_update<RING, PMODE, MMODE, false, false, false>(
                a, nr, i, local, tl, slot, a.k, stim_on, prow, rv, ext, Noise{false, 0.0f}, Params{}, s_new, xs);
        mask = __ballot(spiked);
        if (mask) {
            // queue this wave's spiking rows: lane offsets in the wave's
            // concatenated list, then one reservation for the wave
            const uint32_t len = spiked ? row_hi - row_lo : 0u;
            uint32_t excl = 0, total = 0;
            for (uint64_t m = mask; m;) {
                const uint32_t b = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t lb = __builtin_amdgcn_readlane(len, b);
                if (b < lane) excl += lb;
                total += lb;
            }
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(q_ctr, ((unsigned long long)__popcll(mask) << 32) | total);
            const uint32_t bq = uniform((uint32_t)(base >> 32)), be = uniform((uint32_t)base);
            if (spiked) {
                const uint32_t q = bq + wave_rank(mask);
                q_lo[q] = row_lo;
                q_off[q] = be + excl;
            }
            if (a.xt) {
                for (uint64_t m = mask; m;) {
                    const uint32_t b = (uint32_t)__builtin_ctzll(m);
                    m &= m - 1;
                    push_custom<RING>(a, wave * kWave + b, slot, lane);
                }
            }
        }
        // this step's stores (the pushes below never touch slot t)
        if (act) {
            store_wt(&ring[ri], (RT)0);
            store_wt((float2 *)&a.x[i], xs);
            store_wt((float *)&a.s[i], s_new);
            if (spiked) spike_marks<false>(a, nr, i, t, tl);
        }
        if (lane == 0) spike_words(a, wave, t, rast_origin, mask);
    }
    __syncthreads();

    // ---- the workgroup's deliveries through the bins
    const uint64_t qc = *q_ctr;
    const uint32_t Q = (uint32_t)(qc >> 32), T = (uint32_t)qc;
    if (T == 0) return;            // no spike in the workgroup: nothing to bin (uniform)
    // the table is cleared only in workgroups that use it
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) {
        bkey[k] = kBinEmpty;
        bval[k] = (RT)0;
    }
    __syncthreads();
    {
        using W = RT;
        const W *w = (const W *)a.w;
        const uint32_t shift = 32u - (uint32_t)__builtin_ctz(B);
        for (uint32_t e = threadIdx.x; e < T; e += blockDim.x) {
            // owning row: the last queue entry whose offset is <= e
            uint32_t lo = 0, hi = Q;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (q_off[mid] <= e) lo = mid;
                else hi = mid;
            }
            const uint32_t j = q_lo[lo] + (e - q_off[lo]);
            const uint32_t sy = a.syn[j];
            const W wv = w[j];
            const uint32_t d16 = PACKED ? 0u : a.delay16[j];
            const uint32_t word = (uint32_t)ring_index<PACKED>(a, sy, d16, slot);
            uint32_t h = (word * 0x9E3779B1u) >> shift;
            bool placed = false;
            for (uint32_t p = 0; p < kBinProbes; ++p) {
                const uint32_t prev = atomicCAS(&bkey[h], kBinEmpty, word);
                if (prev == kBinEmpty || prev == word) {
                    atomicAdd(&bval[h], wv);
                    placed = true;
                    break;
                }
                h = (h + 1u) & (B - 1u);
            }
            if (!placed) ring_add_word<RING, PACKED>(a, word, wv);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) {
        const uint32_t word = bkey[k];
        if (word != kBinEmpty) ring_add_word<RING, PACKED>(a, word, bval[k]);
    }
}

// LDS bytes of k_step_binned with `entries` bins and workgroups of `block`.
uint32_t binned_lds_bytes(uint32_t entries, uint32_t block) { return entries * 8u + block * 8u + 8u; }

// ---------------------------------------------------------------- pair
// Two steps per launch, t0 = t and t1 = t + 1.  Between the two steps the
// only dependency that crosses waves is a delay-1 synapse: a spike at t0
// arrives at t1.  Those arrivals are not pushed; the target's wave pulls
// them.  Each wave publishes its t0 spike word (two 64-bit words, each
// holding a 32-bit half and the tag t0+1, stored whole), then reads the
// words of the sources in its delay-1 in-table (source, target lane,
// weight) until they carry the tag, and adds the weights of the sources
// that fired -- a point-to-point wait on the waves that feed it, no grid
// barrier.  Every other delivery of t0 is pushed as in k_step; the ring
// has one slot more than the deepest delay needs, so the pushes of this
// launch (slots t0+2 .. t0+1+maxdelay) never touch slots t0 and t0+1,
// which the launch reads and clears.  (v, u) and s stay in registers
// between the two steps.  With the int ring the pulled weights are the
// same integers the ring would have summed, so the result is the one-step
// result bit for bit.
//
// The wait needs every source wave to run while its consumers wait, so the
// launcher uses this kernel only when the whole grid fits on the device at
// once; the wait is bounded anyway: past a.pair_spins polls (default
// kPairSpinsDefault) a wave raises
// pair_abort (the host reports the launch as failed) and moves on, and
// every other waiting wave sees the flag and stops waiting.

// Step t0 of the launch's first step for one delay-1 source neuron, when
// its wave has not published its word in time (a wave that is not resident
// at all, when another process holds compute units for longer than the
// wait): the source's update recomputed from its inputs, which the launch
// does not change before every wave has published -- x, s and the clear of
// ring slot t0 are the launch's last stores, issued after the wave's own
// publication has completed (see pair_steps) -- and ring slot t0 receives
// no deliveries during the launch; the caller polls the word again after
// the inputs were read, so a source that started meanwhile cannot have
// overwritten them unnoticed.  The same arithmetic as the source's own
// update (same noise key, parameter row and order), so the pulled sum, and
// with the int ring the whole raster, stay exact.  Nothing is written for
// the source.
template <int RING, int PMODE, int MMODE, bool LEAN>
__device__ __forceinline__ bool source_fired(const StepArgs &a, const NetRec &nr, uint32_t src, uint64_t tl0,
                                          uint32_t slot0, uint32_t k, bool stim0, const uint8_t *prow_p) {
    using RT = typename Ring<RING>::T;
    const uint32_t local = src - nr.base;
    if (!(local < nr.n && nr.active && (nr.t_len == 0 || tl0 < nr.t_len))) return false;
    const RT rv = ((const RT *)a.ring)[(size_t)slot0 * a.n_total + src];
    float s = a.s[src];
    float2 xs = a.x[src];
    const float ext = a.ext ? a.ext[src] : 0.0f;
    const uint32_t prow = (PMODE != 1 && prow_p) ? prow_p[src] : 0u;
    return neuron_update<RING, PMODE, MMODE, false, LEAN, false>(a, nr, src, local, tl0, slot0, k, stim0, prow,
                                                                 rv, ext, Noise{false, 0.0f}, Params{}, s, xs);
}

// One 64-entry round of the wave's delay-1 in-table: wait for each entry's
// source word (tag `tag`), then add the weight of every source that fired.
// Past pair_spins polls (or once another wave has given up) the entries
// still unpublished are recomputed (source_fired) instead of waited for,
// and the give-up is noted in the abort flag (a count for the host, not an
// error).
template <int RING, int PMODE, int MMODE, bool LEAN, typename RT>
__device__ __forceinline__ void pull_d1(const StepArgs &a, bool has, uint32_t ent, RT w, uint32_t tag,
                                        RT *acc_w, uint32_t lane, const NetRec &nr, uint64_t tl0,
                                        uint32_t slot0, uint32_t k, bool stim0, const uint8_t *prow_p) {
    const uint32_t src = ent & 0x03FFFFFFu;
    const uint64_t *pw = a.pub + 2u * (src >> 6) + ((src >> 5) & 1u);
    uint64_t v = 0;
    bool fired = false;
    for (uint32_t spins = 0;; ++spins) {
        if (has) v = __hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = !has || (uint32_t)(v >> 32) == tag;
        if (__ballot(!ok) == 0ull) break;
        if (__builtin_expect(spins >= a.pair_spins ||
                                 ((spins & 63u) == 63u &&
                                  __hip_atomic_load(a.pair_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)),
                             0)) {
            if (lane == 0) raise_abort(a);
            if (!ok) {
                fired = source_fired<RING, PMODE, MMODE, LEAN>(a, nr, src, tl0, slot0, k, stim0, prow_p);
                // Poll once more after the inputs were read: still
                // unpublished, the source had not stored over them when
                // they were read (its stores follow its published word);
                // published meanwhile, its word is used instead.
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                v = __hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((uint32_t)(v >> 32) == tag) fired = ((uint32_t)v >> (src & 31u)) & 1u;
            } else {
                fired = ((uint32_t)v >> (src & 31u)) & 1u;
            }
            if (has && fired) atomicAdd(&acc_w[ent >> 26], w);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if (has && (((uint32_t)v >> (src & 31u)) & 1u)) atomicAdd(&acc_w[ent >> 26], w);
}

// The values the first loads of a launch need come as separate leading
// kernel arguments, which the build preloads into scalar registers
// (-amdgpu-kernarg-preload-count): the state and ring loads issue at once,
// without waiting for a load of the argument block (whose other fields are
// needed only after those loads return).
template <int RING, int PMODE, int MMODE, bool PACKED, bool LEAN, bool SPEC>
__device__ __forceinline__ void pair_steps(void *ring_p, float *s_p, float2 *x_p, uint32_t slots,
                                           uint32_t n_total, uint32_t words_wpb, uint32_t k,
                                           const uint8_t *prow_p, const StepArgs &a) {
    using RT = typename Ring<RING>::T;
    __shared__ RT acc[kMaxBlock];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t slot0 = slots & 0xFFFFu, slot1 = slots >> 16;
    const uint32_t words = words_wpb & kWordsMask, wpb = (words_wpb >> kWpbShift) + 1u;
    // waves per workgroup from the arguments (blockDim is a load of the
    // dispatch's hidden arguments, which would delay the first loads)
    const uint32_t wave = blockIdx.x * wpb + uniform(threadIdx.x >> 6);
    if (wave >= words) return;
    const uint32_t i = wave * kWave + lane;
    // ---- every load of the launch up front (the state first, then the
    // network record's dependent loads)
    RT *ring = (RT *)ring_p;
    const size_t r0 = (size_t)slot0 * n_total + i, r1 = (size_t)slot1 * n_total + i;
    const RT rv0 = ring[r0];
    const RT rv1 = ring[r1];
    float s = s_p[i];
    float2 xs = x_p[i];
    uint32_t prow = 0;
    if (PMODE != 1) {
        // parameter-row index, loaded without a branch (where there is no
        // index array a lane reads a byte of its own state and ignores it):
        // a load inside a branch is waited for where the branch ends, and
        // with it every load issued before it
        const uint8_t *pb = prow_p ? prow_p + i : (const uint8_t *)(x_p + i);
        const uint32_t r = *pb;
        prow = prow_p ? r : 0u;
    }
    NetRec nr;
    uint32_t net = 0;
    if (LEAN || a.single_net) {
        nr = a.net0;
    } else {
        net = sload(a.wave_net + wave);
        nr = sload(a.nets + net);
    }
    const float ext = a.ext ? a.ext[i] : 0.0f;
    const uint32_t row_lo = a.row_lo[i], row_hi = a.row_hi[i];
    const uint64_t t0 = sload(a.t_dev) + k;
    const uint64_t rast_origin = a.rast_bits ? sload<uint64_t>(a.rast_ctl + 1) : 0;
    const bool stim0 = !LEAN && a.stim_lo && sload(a.stim_nz + (size_t)k * a.stim_groups + net) != 0u;
    const bool stim1 = !LEAN && a.stim_lo && sload(a.stim_nz + (size_t)(k + 1) * a.stim_groups + net) != 0u;
    // this wave's delay-1 in-table; the first 64 entries are loaded here.
    // Lanes past the table's end load entry 0 (the table always holds one)
    // and never use it: with no branch around the loads, the compiler knows
    // the order of every load in flight and waits for each exactly.
    const uint32_t pl_lo = sload(a.pl_ptr + wave), pl_hi = sload(a.pl_hi + wave);
    const RT *pl_w = (const RT *)a.pl_w;
    const bool has0 = pl_lo + lane < pl_hi;
    const uint32_t e0 = has0 ? pl_lo + lane : 0u;
    const uint32_t ent0 = a.pl_ent[e0];
    const RT w0 = pl_w[e0];

    // test hook (nh_test_pair_wait): wave 0 steps late, as a wave that is
    // not resident would, so the waves it feeds give up waiting first;
    // bounded (1 ms on the 100 MHz real-time counter)
    if (wave == 0u && a.pair_mute) {
        const uint64_t t_start = __builtin_amdgcn_s_memrealtime();
        while (__builtin_amdgcn_s_memrealtime() - t_start < 100000ull) __builtin_amdgcn_s_sleep(127);
    }

    // ---- step t0
    const uint32_t local = i - nr.base;
    const uint64_t tl0 = t0 - (uint64_t)nr.t0;
    const uint64_t t1 = t0 + 1u, tl1 = tl0 + 1u;
    const bool act0 = local < nr.n && nr.active && (nr.t_len == 0 || tl0 < nr.t_len);
    const bool act1 = local < nr.n && nr.active && (nr.t_len == 0 || tl1 < nr.t_len);
    // The noise of both steps depends only on (neuron, step, seed): hashed
    // for every lane here, while the state loads are in flight (a lane that
    // does not step ignores its draws).
    const Noise nz0 = noise_terms(nr, local, tl0);
    const Noise nz1 = noise_terms(nr, local, tl1);
    bool sp0 = false;
    // the parameter row serves both steps
    const Params pp = select_params<PMODE, MMODE, true>(a, i, prow);
    if (act0)
        sp0 = neuron_update<RING, PMODE, MMODE, true, LEAN, true>(a, nr, i, local, tl0, slot0, k, stim0, prow,
                                                                 rv0, ext, nz0, pp, s, xs);
    const uint64_t m0 = __ballot(sp0);
    ready(row_lo);
    ready(row_hi);
    const uint32_t tag = (uint32_t)t0 + 1u;   // never 0 before 2^32 steps: 0 is "unwritten"
    if (lane < 2u) {
        const uint32_t half = lane ? (uint32_t)(m0 >> 32) : (uint32_t)m0;
        __hip_atomic_store(&a.pub[2u * wave + lane], ((uint64_t)tag << 32) | half, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // The push of t0: its first batch's entries are loaded here, so that
    // their round trip runs beside the poll's, and delivered after the
    // pull.  A wave's memory operations complete in order, so atomics (or
    // any access) issued before the poll would hold up its result by their
    // own completion; the spike marks and words wait until after it too.
    ready(ent0);
    PushList pl0{0ull, 0u, 0u};
    PushBatch<RT> pb0;
    if (m0) {
        pl0 = push_first(m0, row_lo, row_hi, lane, pb0.j);
        load_batch<RING, PACKED>(a, pb0);
    }

    // SPEC: step t1 speculatively here, as if no delay-1 arrival
    // came (a pulled sum of zero, the case of nearly every wave), and load
    // the first batch of its push: kept if the wave's pulled sums are all
    // zero (the same operations on the same values: rv1 + 0), redone after
    // the pull otherwise.  The update only reads memory in a paired launch
    // (no custom-tau traces), so a discarded one leaves nothing behind.
    float s_sp = s;
    float2 xs_sp = xs;
    bool sp1_sp = false;
    if (SPEC && act1)
        sp1_sp = neuron_update<RING, PMODE, MMODE, true, LEAN, true>(a, nr, i, local, tl1, slot1, k + 1u,
                                                                    stim1, prow, rv1 + (RT)0, ext, nz1, pp,
                                                                    s_sp, xs_sp);
    const uint64_t m1_sp = SPEC ? __ballot(sp1_sp) : 0ull;
    PushList pl1{0ull, 0u, 0u};
    PushBatch<RT> pb1;
    if (SPEC && m1_sp) {
        pl1 = push_first(m1_sp, row_lo, row_hi, lane, pb1.j);
        load_batch<RING, PACKED>(a, pb1);
    }

    // ---- pull the delay-1 arrivals of t0 into this wave's neurons
    RT *acc_w = acc + (threadIdx.x & ~(uint32_t)(kWave - 1));
    acc_w[lane] = (RT)0;
    // (one call site: the round's give-up path holds a whole neuron update)
    bool has = has0;
    uint32_t ent = ent0;
    RT w = w0;
    for (uint32_t b = pl_lo;;) {
        pull_d1<RING, PMODE, MMODE, LEAN>(a, has, ent, w, tag, acc_w, lane, nr, tl0, slot0, k, stim0, prow_p);
        b += kWave;
        if (b >= pl_hi) break;
        has = b + lane < pl_hi;
        const uint32_t e = has ? b + lane : 0u;
        ent = a.pl_ent[e];
        w = pl_w[e];
        ready(ent);
    }
    if (m0) {
        send_batch<RING, PACKED, true>(a, pb0, slot0);
        push_list_from<RING, PACKED, true, true>(a, pl0, kPushUnroll * kWave, row_lo, row_hi, slot0, lane);
    }
    if (sp0) spike_marks<LEAN>(a, nr, i, t0, tl0);
    if (lane == 0) spike_words(a, wave, t0, rast_origin, m0);
    __builtin_amdgcn_wave_barrier();
    const RT pulled = acc_w[lane];

    // ---- step t1
    bool sp1 = false;
    uint64_t m1;
    if (SPEC && __ballot(pulled != (RT)0) == 0ull) {
        s = s_sp;
        xs = xs_sp;
        sp1 = sp1_sp;
        m1 = m1_sp;
    } else {
        if (act1)
            sp1 = neuron_update<RING, PMODE, MMODE, true, LEAN, true>(a, nr, i, local, tl1, slot1, k + 1u,
                                                                     stim1, prow, rv1 + pulled, ext, nz1,
                                                                     pp, s, xs);
        m1 = __ballot(sp1);
    }
    if (m1) {
        if (SPEC && m1 == m1_sp) {
            send_batch<RING, PACKED, false>(a, pb1, slot1);
            push_list_from<RING, PACKED, false, true>(a, pl1, kPushUnroll * kWave, row_lo, row_hi, slot1, lane);
        } else {
            push_wave<RING, PACKED, false, true>(a, m1, row_lo, row_hi, slot1, lane);
        }
    }

    // ---- stores of both steps.  They overwrite this launch's inputs, which
    // a waiting wave may read to recompute this wave's step t0
    // (source_fired) while this wave's word is unpublished to it: so they
    // are issued only after the publication has completed.  A wave that
    // polled has waited for all its earlier memory operations already (the
    // poll's wait); one with an empty in-table waits here.
    if (pl_lo == pl_hi) __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    if (act0) store_wt(&ring[r0], (RT)0);
    if (act1) store_wt(&ring[r1], (RT)0);
    if (act0 || act1) {
        store_wt(&x_p[i], xs);
        store_wt(&s_p[i], s);
    }
    if (sp1) spike_marks<LEAN>(a, nr, i, t1, tl1);
    if (lane == 0) spike_words(a, wave, t1, rast_origin, m1);
}

template <int RING, int PMODE, int MMODE, bool PACKED, bool LEAN>
__global__ __launch_bounds__(kMaxBlock) void k_pair(void *ring, float *s, float2 *x, uint32_t slots,
                                                    uint32_t n_total, uint32_t words_wpb, uint32_t k,
                                                    const uint8_t *prow, ArgBlock<LEAN> av) {
    pair_steps<RING, PMODE, MMODE, PACKED, LEAN, LEAN>(ring, s, x, slots, n_total, words_wpb, k, prow,
                                                       arg_block(av));
}

// The same kernel without the speculative second step, held to the
// register budget of 8 waves per SIMD (the default build's 73 vector
// registers allow 6): one network, used only where k_pair's grid does not
// fit on the device at once but this one's does (measured when it was added:
// 250k neurons 5.31 -> 3.99 us per step against one step per launch; at 100k
// the spills it took cost 0.14 us per step, so k_pair runs there).
template <int RING, int PMODE, int MMODE, bool PACKED>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_pair8(
    void *ring, float *s, float2 *x, uint32_t slots, uint32_t n_total, uint32_t words_wpb, uint32_t k,
    const uint8_t *prow, ArgBlock<true> av) {
    pair_steps<RING, PMODE, MMODE, PACKED, true, false>(ring, s, x, slots, n_total, words_wpb, k, prow,
                                                        arg_block(av));
}

// ---------------------------------------------------------------- persistent
// Many steps per launch (run_steps) where the whole grid is resident at
// once: (v, u), s, the parameter row, the row bounds and the wave's delay-1
// in-table stay in registers for the whole launch.  Per step t:
//  1. wait until every workgroup has finished step t-2, pushes complete (per
//     group of workgroups a done tag, published by the group's last
//     workgroup to arrive).  Deliveries of delay >= 2 into slot t come from
//     steps <= t-2.  The first wave of a workgroup to see the tags notes
//     them in LDS for the others.
//  2. read (and, after the update, clear) ring slot t with agent-scope
//     loads and stores: during the launch the slot is written by atomics
//     from every XCD, and the XCD's L2 may hold an old copy of the line,
//     which these bypass.
//  3. pull the delay-1 arrivals of step t-1 from the published spike words
//     of the in-table sources (as k_pair does for its second step).
//  4. update, ballot, publish the spike word of step t (three buffers),
//     push every delivery except the delay-1 ones (all of them at the
//     launch's last step: the next launch reads those from the ring).
//  5. wait for this wave's memory operations (ring clear, pushes) to
//     complete, then arrive: an LDS count per workgroup, whose last wave
//     adds to its group's counter; the group's last workgroup resets the
//     counter and publishes the done tag t + 1.
// Waves are at most one step apart: a wave at step t has passed the wait
// for t-2, so every wave is at t-1 or later, and a wave at t+1 needs t-1
// finished everywhere.  Hence the ring's two slots beyond the deepest delay
// (as for k_pair: a push of step t reaches slot t+maxd, which a wave still
// at t-1 may be reading), three buffers of published words and four counter
// residues.
// No fences: every access to memory shared within the launch (ring words,
// published words, counters, done tags) is an agent-scope atomic access,
// coherent across XCDs by itself, and a wave waits for its own accesses to
// complete (s_waitcnt vmcnt(0)) before it signals; a waiter issues the
// loads that depend on a signal only after it has read the signal.  (An
// agent-scope release/acquire fence would add an L2 write-back and
// invalidate for non-atomic data, of which there is none here.)
// Every wait is bounded as in k_pair (pair_spins polls, then the abort
// flag, which every other wait also checks): the launch always ends.
// Measured on the 100k network: 4.06 us per step over 2000-step runs and
// 4.2 us per step over 20-step runs, against 2.87 and 3.34 for k_pair at
// the time (2.27 and 3.3 since) --
// each step here chains about five memory round trips across XCDs (gate,
// ring read, pull, push, completion) where a pair of steps chains about
// four plus a launch -- so it runs only when asked for (step shape 3).

// Ring word access at agent scope, as 32-bit words (float or int32 ring).
template <typename RT>
__device__ __forceinline__ RT ring_load_agent(const RT *p) {
    const uint32_t b = __hip_atomic_load((const uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(RT, b);
}
template <typename RT>
__device__ __forceinline__ void ring_clear_agent(RT *p) {
    __hip_atomic_store((uint32_t *)p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded-wait bookkeeping: true when the wave must stop waiting (limit
// reached or another wave gave up), after raising the abort flag.
__device__ __forceinline__ bool wait_expired(const StepArgs &a, uint32_t spins, uint32_t lane) {
    if (spins >= a.pair_spins ||
        ((spins & 63u) == 63u && __hip_atomic_load(a.pair_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        if (lane == 0) raise_abort(a);
        return true;
    }
    return false;
}

// Wait until step `st` is finished everywhere: every group's done tag for
// residue st % 4 reads st + 1 (lanes g < n_groups poll one group each).
// *seen (LDS) holds the latest step + 1 a wave of the workgroup has seen
// finished; waves check it before polling memory and note what they saw.
__device__ __forceinline__ bool wait_done(const StepArgs &a, uint64_t st, uint32_t *seen, uint32_t lane) {
    const uint32_t tag = (uint32_t)st + 1u;
    if ((int32_t)(*(volatile uint32_t *)seen - tag) >= 0) return true;
    const bool mine = lane < a.n_groups;
    const uint32_t *pw = a.done + (size_t)(st & 3u) * a.n_groups + (mine ? lane : 0u);
    for (uint32_t spins = 0;; ++spins) {
        uint32_t v = tag;
        if (mine) v = __hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__ballot(v != tag) == 0ull) {
            if (lane == 0) atomicMax(seen, tag);
            return true;
        }
        if (wait_expired(a, spins, lane)) return false;
        __builtin_amdgcn_s_sleep(1);
        if ((int32_t)(*(volatile uint32_t *)seen - tag) >= 0) return true;
    }
}

// One 64-entry round of a wave's delay-1 in-table from the published words
// `pub` (tag `tag`); false when the wait ran out or was aborted.
template <typename RT>
__device__ __forceinline__ bool pull_words(const StepArgs &a, const uint64_t *pub, bool has, uint32_t ent,
                                           RT w, uint32_t tag, RT *acc_w, uint32_t lane) {
    const uint32_t src = ent & 0x03FFFFFFu;
    const uint64_t *pw = pub + 2u * (src >> 6) + ((src >> 5) & 1u);
    uint64_t v = 0;
    for (uint32_t spins = 0;; ++spins) {
        if (has) v = __hip_atomic_load(pw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool ok = !has || (uint32_t)(v >> 32) == tag;
        if (__ballot(!ok) == 0ull) break;
        if (wait_expired(a, spins, lane)) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    if (has && (((uint32_t)v >> (src & 31u)) & 1u)) atomicAdd(&acc_w[ent >> 26], w);
    return true;
}

template <int RING, int PMODE, int MMODE, bool PACKED>
__global__ __launch_bounds__(kMaxBlock) void k_run(StepArgs a) {
    using RT = typename Ring<RING>::T;
    __shared__ RT acc[kMaxBlock];
    __shared__ uint32_t arrive[4];
    __shared__ uint32_t seen;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t nwb = blockDim.x >> 6;
    const uint32_t wave = blockIdx.x * nwb + uniform(threadIdx.x >> 6);
    if (threadIdx.x < 4u) arrive[threadIdx.x] = 0u;
    if (threadIdx.x == 0u) seen = 0u;
    __syncthreads();
    if (wave >= a.words) return;
    // waves of this workgroup that take part (the last workgroup may be
    // partial), and the workgroups of its group
    const uint32_t nvalid = min(nwb, a.words - blockIdx.x * nwb);
    const uint32_t grp = blockIdx.x % a.n_groups;
    const uint32_t grp_wgs = gridDim.x / a.n_groups + (grp < gridDim.x % a.n_groups ? 1u : 0u);
    const uint32_t i = wave * kWave + lane;
    const NetRec nr = a.net0;
    RT *ring = (RT *)a.ring;

    // ---- state and tables for the whole launch
    float s = a.s[i];
    float2 xs = a.x[i];
    const float ext = a.ext ? a.ext[i] : 0.0f;
    const uint32_t row_lo = a.row_lo[i], row_hi = a.row_hi[i];
    uint32_t prow = 0;
    if (PMODE != 1 && (PMODE == 2 || a.n_krows > 1)) prow = a.prow[i];
    const uint64_t t0 = sload(a.t_dev) + a.k;
    const uint64_t rast_origin = a.rast_bits ? sload<uint64_t>(a.rast_ctl + 1) : 0;
    const uint32_t pl_lo = sload(a.pl_ptr + wave), pl_hi = sload(a.pl_hi + wave);
    const RT *pl_w = (const RT *)a.pl_w;
    const bool has0 = pl_lo + lane < pl_hi;
    uint32_t ent0 = 0;
    RT w0 = (RT)0;
    if (has0) {
        ent0 = a.pl_ent[pl_lo + lane];
        w0 = pl_w[pl_lo + lane];
    }
    const uint32_t local = i - nr.base;
    const bool act = local < nr.n;
    const Params pp = select_params<PMODE, MMODE, true>(a, i, prow);
    RT *acc_w = acc + (threadIdx.x & ~(uint32_t)(kWave - 1));
    const uint32_t S = a.run_steps;
    const size_t pub_stride = 2u * (size_t)a.words;
    uint32_t slot = a.slot;

    for (uint32_t k = 0; k < S; ++k) {
        const uint64_t t = t0 + k;
        const uint64_t tl = t - (uint64_t)nr.t0;
        // ---- 1. every push into slot t is complete
        if (k >= 2u && !wait_done(a, t - 2u, &seen, lane)) break;
        // ---- 2. ring slot t
        RT *rp = ring + (size_t)slot * a.n_total + i;
        RT rv = ring_load_agent(rp);
        // ---- 3. delay-1 arrivals of step t-1 (published in this launch)
        if (k >= 1u) {
            acc_w[lane] = (RT)0;
            const uint64_t *pubp = a.pub3 + (size_t)((t - 1u) % kPubDepth) * pub_stride;
            const uint32_t tag = (uint32_t)t;       // step t-1 publishes (t-1)+1
            bool ok = pull_words<RT>(a, pubp, has0, ent0, w0, tag, acc_w, lane);
            for (uint32_t b = pl_lo + kWave; ok && b < pl_hi; b += kWave) {
                const bool has = b + lane < pl_hi;
                uint32_t ent = 0;
                RT w = (RT)0;
                if (has) {
                    ent = a.pl_ent[b + lane];
                    w = pl_w[b + lane];
                }
                ok = pull_words<RT>(a, pubp, has, ent, w, tag, acc_w, lane);
            }
            if (!ok) break;
            __builtin_amdgcn_wave_barrier();
            rv = rv + acc_w[lane];
        }
        // ---- 4. update, publish, push
        bool sp = false;
        if (act)
            sp = neuron_update<RING, PMODE, MMODE, false, true, true>(a, nr, i, local, tl, slot, 0u, false, prow,
                                                                     rv, ext, Noise{false, 0.0f}, pp, s, xs);
        const uint64_t m = __ballot(sp);
        ring_clear_agent(rp);       // after its value is used: no store overtakes the load
        const bool last = k + 1u == S;
        if (!last && lane < 2u && !(a.pair_mute && wave == 0u)) {
            const uint32_t half = lane ? (uint32_t)(m >> 32) : (uint32_t)m;
            __hip_atomic_store(a.pub3 + (size_t)(t % kPubDepth) * pub_stride + 2u * wave + lane,
                               ((uint64_t)((uint32_t)t + 1u) << 32) | half, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        if (m) {
            if (last) push_wave<RING, PACKED, false>(a, m, row_lo, row_hi, slot, lane);
            else push_wave<RING, PACKED, true>(a, m, row_lo, row_hi, slot, lane);
        }
        if (sp) spike_marks<true>(a, nr, i, t, tl);
        if (lane == 0) spike_words(a, wave, t, rast_origin, m);
        slot = slot + 1u == a.D ? 0u : slot + 1u;
        // ---- 5. arrive (only steps a later step of the launch waits for)
        if (k + 3u > S) continue;
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (lane == 0) {
            const uint32_t r = (uint32_t)t & 3u;
            const uint32_t before = atomicAdd(&arrive[r], 1u);
            if (before + 1u == nvalid) {
                arrive[r] = 0u;
                uint32_t *gc = a.gctr + (size_t)r * a.n_groups + grp;
                const uint32_t g_before = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (g_before + 1u == grp_wgs) {
                    __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                    __hip_atomic_store(a.done + (size_t)r * a.n_groups + grp, (uint32_t)t + 1u,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    if (act) {
        store_wt((float2 *)&a.x[i], xs);
        store_wt((float *)&a.s[i], s);
    }
}

template <int RING, int PMODE, int MMODE>
static const void *run_fn(bool packed) {
    return packed ? (const void *)k_run<RING, PMODE, MMODE, true> : (const void *)k_run<RING, PMODE, MMODE, false>;
}

static const void *run_kernel(int ring, int pmode, int mmode, bool packed) {
#define NH_RUN_M(R, P)                                                                            \
    switch (mmode) {                                                                              \
    case 0: return run_fn<R, P, 0>(packed);                                                       \
    case 1: return run_fn<R, P, 1>(packed);                                                       \
    default: return run_fn<R, P, 2>(packed);                                                      \
    }
#define NH_RUN_P(R)                                                                               \
    switch (pmode) {                                                                              \
    case 0: NH_RUN_M(R, 0)                                                                        \
    case 1: NH_RUN_M(R, 1)                                                                        \
    default: NH_RUN_M(R, 2)                                                                       \
    }
    if (ring == 0) { NH_RUN_P(0) }
    NH_RUN_P(1)
#undef NH_RUN_P
#undef NH_RUN_M
}

void launch_run(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool packed, hipStream_t st) {
    void *args[] = {(void *)&a};
    (void)hipLaunchKernel(run_kernel(ring, pmode, mmode, packed), dim3(blocks), dim3(a.block), args, 0, st);
}

// Waves of the persistent kernel the device holds at once with workgroups
// of `block` threads (0 when the query fails).
uint32_t run_resident_waves(int ring, int pmode, int mmode, bool packed, uint32_t block) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, run_kernel(ring, pmode, mmode, packed), (int)block,
                                                     0) != hipSuccess)
        return 0;
    return (uint32_t)cus * (uint32_t)per_cu * (block / 64u);
}

// ---------------------------------------------------------------- scatter
// Two-kernel shape: one wave per spike of step t, all synapses of the row.
template <int RING, bool PACKED>
__global__ __launch_bounds__(kMaxBlock) void k_scatter(StepArgs a) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + uniform(threadIdx.x >> 6);
    const uint64_t t = sload(a.t_dev) + a.k;
    const uint32_t L = (uint32_t)(t % 3);
    const uint32_t c = a.cnt[L];
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t q = wave; q < c; q += nwaves) {
        const uint32_t src = uniform(a.spk[(size_t)L * a.n_total + q]);
        const uint32_t lo = a.row_lo[src], hi = a.row_hi[src];
        push_row<RING, PACKED>(a, lo, hi, a.slot, lane);
        if (a.xt) push_custom<RING>(a, src, a.slot, lane);
    }
}

// Watched-neuron state rows (sim/src/kernels.rs:141-147 moved out of the
// per-neuron loop): one thread per watched neuron after the step.
__global__ void k_watch(StepArgs a) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n_watch) return;
    const uint64_t t = sload(a.t_dev) + a.k;
    a.xrec[(size_t)(t % a.xrec_depth) * a.n_watch + w] = a.x[a.watch[w]];
}

// Per-neuron spike counts over the record ring (sim/src/cuda.rs:725-763
// counts on the host after copying the whole ring): arena neurons
// [base, base + n), base a multiple of 64, over the `rows` steps before step
// t; ring of R rows, one 64-bit word per wave and step.  One thread per
// neuron; a wave's word of a row is one wave-uniform load, eight rows in
// flight.  Only n counts leave the device.
__global__ void k_count_rows(const uint64_t *bits, uint32_t R, uint32_t words, uint32_t base,
                             uint32_t n, uint64_t t, uint32_t rows, uint32_t *out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    // whole waves past the range leave (their word index would pass the row)
    if (uniform(j & ~63u) >= n) return;
    const uint32_t i = base + j;
    const uint32_t wi = uniform(i >> 6);
    const uint32_t bit = i & 63u;
    uint32_t c = 0;
    uint32_t row = (uint32_t)((t - 1u) % R);
    uint32_t b = 0;
    for (; b + 8u <= rows; b += 8u) {
        uint64_t w[8];
#pragma unroll
        for (uint32_t q = 0; q < 8u; ++q) {
            w[q] = bits[(size_t)row * words + wi];
            row = row == 0u ? R - 1u : row - 1u;
        }
#pragma unroll
        for (uint32_t q = 0; q < 8u; ++q) c += (uint32_t)(w[q] >> bit) & 1u;
    }
    for (; b < rows; ++b) {
        c += (uint32_t)(bits[(size_t)row * words + wi] >> bit) & 1u;
        row = row == 0u ? R - 1u : row - 1u;
    }
    if (j < n) out[j] = c;
}

// Count scratch for callers that cannot allocate device memory
// (nh_spike_counts_noalloc): a file-scope device array, 16 MiB, of which
// every GPU the library runs on holds its own copy until the process ends.
// On the host `scratch` is a stand-in, not a device address: host code uses
// only the address scratch_on_device() returns, looked up with
// hipGetSymbolAddress once per device with that device current.
constexpr uint32_t kScratchWordsC = 1u << 22;
static __device__ __align__(256) uint32_t scratch[kScratchWordsC];
extern const uint32_t kScratchWords = kScratchWordsC;

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint32_t *g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_on_device(int dev, uint32_t **ptr) {
    *ptr = nullptr;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = (uint32_t *)p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

void launch_count_rows(const uint64_t *bits, uint32_t R, uint32_t words, uint32_t base, uint32_t n,
                       uint64_t t, uint32_t rows, uint32_t *out, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_count_rows, dim3((n + 255) / 256), dim3(256), 0, st, bits, R, words, base, n, t,
                       rows, out);
}

// Served arena: a newly admitted range starts with empty ring columns,
// zero current and counters, and no last spike.
__global__ void k_clear_range(StepArgs a, uint32_t base, uint32_t len) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    const uint32_t i = base + j;
    uint32_t *ring = (uint32_t *)a.ring;
    for (uint32_t d = 0; d < a.D; ++d) ring[(size_t)d * a.n_total + i] = 0u;
    a.s[i] = 0.0f;
    a.pn_total[i] = 0u;
    a.last_spike[i] = kNoSpike;
    if (a.wcnt)
        for (uint32_t w = 0; w < a.n_win; ++w) a.wcnt[(size_t)w * a.n_total + i] = 0u;
}

void launch_clear_range(const StepArgs &a, uint32_t base, uint32_t len, hipStream_t st) {
    if (!len) return;
    hipLaunchKernelGGL(k_clear_range, dim3((len + 255) / 256), dim3(256), 0, st, a, base, len);
}

// Plain copy of n words of T (byte-counter calibration; T = uint32_t gives
// the step kernel's 4-byte lanes, uint4 the widest loads).
template <typename T>
__global__ void k_copy(const T *__restrict__ src, T *__restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

void launch_copy(const void *src, void *dst, uint64_t bytes, uint32_t lane_bytes, hipStream_t st) {
    if (lane_bytes == 16)
        hipLaunchKernelGGL(k_copy<uint4>, dim3(4096), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst,
                           bytes / 16);
    else
        hipLaunchKernelGGL(k_copy<uint32_t>, dim3(8192), dim3(256), 0, st, (const uint32_t *)src,
                           (uint32_t *)dst, bytes / 4);
}

// Empty kernel: the dispatch cost of one graph node (launch-floor probe).
__global__ void k_empty() {}

void launch_empty(uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_empty, dim3(blocks), dim3(256), 0, st);
}

// End of a chunk: advance the device step base.
__global__ void k_advance(uint64_t *t_dev, uint32_t k) { *t_dev += k; }

// ------------------------------------------------------------ dispatch
template <int RING, int PMODE, int MMODE, bool PUSH, bool PACKED, bool LEAN>
static void go(const StepArgs &a, uint32_t blocks, hipStream_t st) {
    // a packed arena's grid is padded to whole bands (waves past the arena
    // return at once)
    if (!LEAN && PUSH && !a.single_net && blocks >= 4u * kBands) blocks = (blocks + kBands - 1u) / kBands * kBands;
    const uint32_t band_len = (!LEAN && PUSH && !a.single_net && blocks % kBands == 0u) ? blocks / kBands : 0u;
    const uint32_t lk = a.slot | a.k << kSlotBits | (a.block / 64u - 1u) << kWpbShift;
    hipLaunchKernelGGL((k_step<RING, PMODE, MMODE, PUSH, PACKED, LEAN>), dim3(blocks), dim3(a.block), 0, st,
                       a.ring, a.s, a.x, a.n_total, lk, band_len, prow_arg(a, PMODE), arg_pass<LEAN>(a));
}

template <int RING, int PMODE, int MMODE>
static void launch_m(const StepArgs &a, uint32_t blocks, bool push, bool packed, bool lean, hipStream_t st) {
    if (!push) go<RING, PMODE, MMODE, false, false, false>(a, blocks, st);
    else if (packed && lean) go<RING, PMODE, MMODE, true, true, true>(a, blocks, st);
    else if (packed) go<RING, PMODE, MMODE, true, true, false>(a, blocks, st);
    else if (lean) go<RING, PMODE, MMODE, true, false, true>(a, blocks, st);
    else go<RING, PMODE, MMODE, true, false, false>(a, blocks, st);
}

template <int RING, int PMODE>
static void launch_p(const StepArgs &a, uint32_t blocks, int mmode, bool push, bool packed, bool lean,
                     hipStream_t st) {
    switch (mmode) {
    case 0: launch_m<RING, PMODE, 0>(a, blocks, push, packed, lean, st); break;
    case 1: launch_m<RING, PMODE, 1>(a, blocks, push, packed, lean, st); break;
    default: launch_m<RING, PMODE, 2>(a, blocks, push, packed, lean, st); break;
    }
}

template <int RING>
static void launch_r(const StepArgs &a, uint32_t blocks, int pmode, int mmode, bool push, bool packed,
                     bool lean, hipStream_t st) {
    switch (pmode) {
    case 0: launch_p<RING, 0>(a, blocks, mmode, push, packed, lean, st); break;
    case 1: launch_p<RING, 1>(a, blocks, mmode, push, packed, lean, st); break;
    default: launch_p<RING, 2>(a, blocks, mmode, push, packed, lean, st); break;
    }
}

template <int RING, int PMODE, int MMODE>
static void go_binned(const StepArgs &a, uint32_t blocks, bool packed, hipStream_t st) {
    if (!a.single_net && blocks >= 4u * kBands) blocks = (blocks + kBands - 1u) / kBands * kBands;
    const uint32_t lds = binned_lds_bytes(a.bin_entries, a.block);
    if (packed)
        hipLaunchKernelGGL((k_step_binned<RING, PMODE, MMODE, true>), dim3(blocks), dim3(a.block), lds, st, a);
    else
        hipLaunchKernelGGL((k_step_binned<RING, PMODE, MMODE, false>), dim3(blocks), dim3(a.block), lds, st, a);
}

template <int RING>
static void binned_r(const StepArgs &a, uint32_t blocks, int pmode, int mmode, bool packed, hipStream_t st) {
#define NH_BIN_M(P)                                                                               \
    switch (mmode) {                                                                              \
    case 0: go_binned<RING, P, 0>(a, blocks, packed, st); break;                                  \
    case 1: go_binned<RING, P, 1>(a, blocks, packed, st); break;                                  \
    default: go_binned<RING, P, 2>(a, blocks, packed, st); break;                                 \
    }
    switch (pmode) {
    case 0: NH_BIN_M(0) break;
    case 1: NH_BIN_M(1) break;
    default: NH_BIN_M(2) break;
    }
#undef NH_BIN_M
}

void launch_step_binned(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool packed,
                        hipStream_t st) {
    if (ring == 0) binned_r<0>(a, blocks, pmode, mmode, packed, st);
    else binned_r<1>(a, blocks, pmode, mmode, packed, st);
}

void launch_step(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool push,
                 bool packed, bool lean, hipStream_t st) {
    if (ring == 0) launch_r<0>(a, blocks, pmode, mmode, push, packed, lean, st);
    else launch_r<1>(a, blocks, pmode, mmode, push, packed, lean, st);
}

template <int RING, int PMODE, int MMODE>
static const void *pair_fn(bool packed, bool lean, bool w8) {
    if (w8) return packed ? (const void *)k_pair8<RING, PMODE, MMODE, true>
                          : (const void *)k_pair8<RING, PMODE, MMODE, false>;
    if (packed) return lean ? (const void *)k_pair<RING, PMODE, MMODE, true, true>
                            : (const void *)k_pair<RING, PMODE, MMODE, true, false>;
    return lean ? (const void *)k_pair<RING, PMODE, MMODE, false, true>
                : (const void *)k_pair<RING, PMODE, MMODE, false, false>;
}

// w8: the 8-waves-per-SIMD build (lean only)
static const void *pair_kernel(int ring, int pmode, int mmode, bool packed, bool lean, bool w8) {
#define NH_PAIR_M(R, P)                                                                           \
    switch (mmode) {                                                                              \
    case 0: return pair_fn<R, P, 0>(packed, lean, w8);                                            \
    case 1: return pair_fn<R, P, 1>(packed, lean, w8);                                            \
    default: return pair_fn<R, P, 2>(packed, lean, w8);                                           \
    }
#define NH_PAIR_P(R)                                                                              \
    switch (pmode) {                                                                              \
    case 0: NH_PAIR_M(R, 0)                                                                       \
    case 1: NH_PAIR_M(R, 1)                                                                       \
    default: NH_PAIR_M(R, 2)                                                                      \
    }
    if (ring == 0) { NH_PAIR_P(0) }
    NH_PAIR_P(1)
#undef NH_PAIR_P
#undef NH_PAIR_M
}

void launch_pair(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool packed,
                 bool lean, bool w8, hipStream_t st) {
    void *ring_p = a.ring;
    float *s_p = a.s;
    float2 *x_p = a.x;
    uint32_t slots = a.slot | (a.slot + 1u == a.D ? 0u : a.slot + 1u) << 16, n_total = a.n_total,
             words_wpb = pack_words_wpb(a);
    const uint8_t *prow = prow_arg(a, pmode);
    uint32_t k = a.k;
    const StepArgs *pa = a.self;
    // the lean builds read the argument block's device copy, the others the
    // block itself (see ArgBlock)
    void *args[] = {&ring_p, &s_p, &x_p, &slots, &n_total, &words_wpb, &k, &prow,
                    lean ? (void *)&pa : (void *)&a};
    (void)hipLaunchKernel(pair_kernel(ring, pmode, mmode, packed, lean, w8 && lean), dim3(blocks),
                          dim3(a.block), args, 0, st);
}

// Waves of the paired-step kernel that the device holds at once with
// workgroups of `block` threads (0 when the query fails).
uint32_t pair_resident_waves(int ring, int pmode, int mmode, bool packed, bool lean, bool w8,
                             uint32_t block) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pair_kernel(ring, pmode, mmode, packed, lean,
                                                                          w8 && lean),
                                                     (int)block, 0) != hipSuccess)
        return 0;
    return (uint32_t)cus * (uint32_t)per_cu * (block / 64u);
}

void launch_scatter(const StepArgs &a, uint32_t blocks, int ring, bool packed, hipStream_t st) {
    const dim3 b(a.block);
    if (ring == 0) {
        if (packed) hipLaunchKernelGGL((k_scatter<0, true>), dim3(blocks), b, 0, st, a);
        else hipLaunchKernelGGL((k_scatter<0, false>), dim3(blocks), b, 0, st, a);
    } else {
        if (packed) hipLaunchKernelGGL((k_scatter<1, true>), dim3(blocks), b, 0, st, a);
        else hipLaunchKernelGGL((k_scatter<1, false>), dim3(blocks), b, 0, st, a);
    }
}

void launch_watch(const StepArgs &a, hipStream_t st) {
    const uint32_t blocks = (a.n_watch + 255) / 256;
    hipLaunchKernelGGL(k_watch, dim3(blocks), dim3(256), 0, st, a);
}

void launch_advance(uint64_t *t_dev, uint32_t k, hipStream_t st) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, t_dev, k);
}

}  // namespace nh
