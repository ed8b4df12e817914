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
//    ring with hardware atomics (lanes stride the delay-sorted row).  No
//    cross-wave dependency exists inside a step, so there is no grid
//    barrier; the kernel boundary orders step t's pushes before step t+1's
//    reads.
//
//  * k_step<PUSH=false> + k_scatter ("two-kernel" shape): the update
//    compacts the spikes of the step into a list; a second kernel scatters
//    one spike per wave.  Kept as the reference shape and for comparison.
//
// The int ring (sim/src/kernels.rs:57-66) sums integers, which commute, so
// both shapes give the CPU raster bit for bit whatever the atomic order.
#include <hip/hip_runtime.h>
#include "device.h"
#include "layout.h"

namespace nh {

template <int RING> struct Ring;
template <> struct Ring<0> {
    using T = float;
    static __device__ __forceinline__ float get(T v, float) { return v; }
    static __device__ __forceinline__ void add(T *p, T v) { unsafeAtomicAdd(p, v); }
};
template <> struct Ring<1> {
    using T = int32_t;
    // sim/src/kernels.rs:60: (float)x / WSCALE
    static __device__ __forceinline__ float get(T v, float scale) { return (float)v / scale; }
    static __device__ __forceinline__ void add(T *p, T v) { atomicAdd(p, v); }
};

__device__ __forceinline__ uint32_t wave_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Push one spike's out-synapse row into the ring; lanes stride the row
// (delay-sorted, so a run of lanes hits one ring row with ascending targets).
template <int RING, bool PACKED>
__device__ __forceinline__ void push_row(const StepArgs &a, uint32_t lo, uint32_t hi,
                                         uint32_t slot0, uint32_t lane) {
    using RT = typename Ring<RING>::T;
    RT *ring = (RT *)a.ring;
    const RT *w = (const RT *)a.w;
    for (uint32_t j = lo + lane; j < hi; j += kWave) {
        const uint32_t sy = a.syn[j];
        uint32_t tgt, d;
        if (PACKED) {
            tgt = sy & 0x07FFFFFFu;
            d = sy >> 27;
        } else {
            tgt = sy;
            d = a.delay16[j];
        }
        uint32_t sl = slot0 + d;
        sl = sl >= a.D ? sl - a.D : sl;
        Ring<RING>::add(ring + (size_t)sl * a.n_total + tgt, w[j]);
    }
}

// Custom-tau synapses of one spike: arrival marks in the per-synapse ring
// (cpu.rs:271-273 pushes the synapse id; here the weight lands in its slot).
template <int RING>
__device__ __forceinline__ void push_custom(const StepArgs &a, uint32_t src, uint32_t slot0,
                                            uint32_t lane) {
    using RT = typename Ring<RING>::T;
    RT *xr = (RT *)a.xring;
    const RT *w = (const RT *)a.xs_w;
    const uint32_t lo = a.xs_lo[src], hi = a.xs_hi[src];
    for (uint32_t e = lo + lane; e < hi; e += kWave) {
        uint32_t sl = slot0 + a.xs_d[e];
        sl = sl >= a.D ? sl - a.D : sl;
        Ring<RING>::add(xr + (size_t)sl * a.n_x + a.xs_k[e], w[e]);
    }
}

// ---------------------------------------------------------------- step
// PMODE: 0 parameter rows in the kernel arguments (<= 4 rows), 1 per-neuron
// rows, 2 global row table.  MMODE: 0 izh, 1 adex, 2 mixed.
// LIST: keep the per-step spike list (two-kernel shape or raster on).
template <int RING, int PMODE, int MMODE, bool PUSH, bool PACKED, bool LIST>
__global__ __launch_bounds__(kMaxBlock) void k_step(StepArgs a) {
    using RT = typename Ring<RING>::T;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + uniform(threadIdx.x >> 6);
    if (wave >= a.words) return;
    const uint32_t i = wave * kWave + lane;
    const uint32_t slot = a.slot;
    const uint32_t net = a.single_net ? 0u : uniform((uint32_t)a.wave_net[wave]);
    const NetRec &nr = a.single_net ? a.net0 : a.nets[net];

    // ---- first batch of independent loads
    RT *ring = (RT *)a.ring;
    const size_t ri = (size_t)slot * a.n_total + i;
    const RT rv = ring[ri];
    const float s_old = a.s[i];
    float2 xs = a.x[i];
    const float ext = a.ext ? a.ext[i] : 0.0f;
    uint32_t row_lo = 0, row_hi = 0;
    if (PUSH) {
        row_lo = a.row_lo[i];
        row_hi = a.row_hi[i];
    }
    uint32_t prow = 0;
    if (PMODE != 1 && (PMODE == 2 || a.n_krows > 1)) prow = a.prow[i];
    const uint64_t t = *a.t_dev + a.k;

    const uint32_t local = i - nr.base;
    const uint64_t tl = t - (uint64_t)nr.t0;
    const bool act = local < nr.n && nr.active && (nr.t_len == 0 || tl < nr.t_len);
    bool spiked = false;
    if (act) {
        ring[ri] = (RT)0;
        const float arr = Ring<RING>::get(rv, nr.int_scale);
        const float sv = s_old + arr;
        float I = sv + ext;
        if (a.xt_ptr) {
            // custom-tau traces into this neuron (cpu.rs:208-228): this
            // step's arrival, use, then decay -- each slot has one target,
            // so its thread owns it
            RT *xr = (RT *)a.xring + (size_t)slot * a.n_x;
            for (uint32_t j = a.xt_ptr[i]; j < a.xt_ptr[i + 1]; ++j) {
                const uint32_t kx = a.xt_idx[j];
                const RT av = xr[kx];
                xr[kx] = (RT)0;
                const float xv = a.xsyn[kx] + Ring<RING>::get(av, nr.int_scale);
                I += xv;
                a.xsyn[kx] = xv * a.xdecay[kx];
            }
        }
        if (a.stim_lo) {
            const float *row = a.stim_win + (size_t)a.k * a.stim_cols + nr.stim_base;
            for (uint32_t j = a.stim_lo[i]; j < a.stim_hi[i]; ++j)
                I += a.stim_w[j] * row[a.stim_input[j]];
        }
        float pr[kMaxParams];
        uint32_t model = (uint32_t)MMODE;
        if (PMODE == 0) {
#pragma unroll
            for (int q = 0; q < kMaxParams; ++q) pr[q] = a.krow[0].p[q];
            model = a.krow[0].model;
#pragma unroll
            for (int r = 1; r < kKargRows; ++r) {
                const bool sel = prow == (uint32_t)r;
#pragma unroll
                for (int q = 0; q < kMaxParams; ++q) pr[q] = sel ? a.krow[r].p[q] : pr[q];
                model = sel ? a.krow[r].model : model;
            }
        } else if (PMODE == 2) {
            const PRow &r = a.ptab[prow];
#pragma unroll
            for (int q = 0; q < kMaxParams; ++q) pr[q] = r.p[q];
            model = r.model;
        } else {
            model = MMODE == 2 ? a.mreg[i] : (uint32_t)MMODE;
            if (MMODE == 0 && a.P == 4) {
                const float4 v = *(const float4 *)(a.p + (size_t)i * 4);
                pr[0] = v.x; pr[1] = v.y; pr[2] = v.z; pr[3] = v.w;
            } else {
                for (uint32_t q = 0; q < a.P && q < (uint32_t)kMaxParams; ++q)
                    pr[q] = a.p[(size_t)i * a.P + q];
            }
        }
        const bool is_adex = MMODE == 1 || (MMODE == 2 && model == 1);
        // sim/src/cpu.rs:240-245: kick * model gain (C for AdEx, adex.rs:189-191)
        const float gain = is_adex ? pr[0] : 1.0f;
        if (nr.pkick > 0.0f && kick_fires(local, tl, nr.seed, nr.pkick)) I += nr.kick * gain;
        if (nr.sd != 0.0f) I += (nr.sd * gauss_draw(local, tl, nr.seed)) * gain;
        if (is_adex)
            spiked = adex_step(pr, xs, I, nr.dt);
        else
            spiked = izh_step(pr[0], pr[1], pr[2], pr[3], xs, I, nr.dt);
        a.x[i] = xs;
        a.s[i] = sv * nr.decay_s;
        if (spiked) {
            a.pn_total[i] += 1;
            a.last_spike[i] = t < 0xFFFFFFFEull ? (uint32_t)t : 0xFFFFFFFEu;
            if (a.wcnt && nr.win_steps) {
                const uint64_t wi = tl / nr.win_steps;
                if (wi < a.n_win) a.wcnt[(size_t)wi * a.n_total + i] += 1;
            }
        }
    }

    // ---- wave64 ballot: the mask IS this wave's word of the spike bitmap
    const uint64_t mask = __ballot(spiked);
    if (a.bits && lane == 0) a.bits[(size_t)(t % a.bits_rows) * a.words + wave] = mask;
    const uint32_t L = (uint32_t)(t % 3);
    if (LIST && blockIdx.x == 0 && threadIdx.x == 0) {
        // spike-list bookkeeping of step t (every value read here is final:
        // written by earlier launches)
        const uint32_t Lp = (L + 2) % 3, Ln = (L + 1) % 3;
        const uint64_t base_t = a.totals[2 + Lp] + a.cnt[Lp];
        a.totals[2 + L] = base_t;
        a.cnt[Ln] = 0;
        if (a.rast_off) {
            const uint64_t rs = t - a.rast_ctl[1];
            if (rs < a.rast_cap_steps) a.rast_off[rs] = (uint32_t)(base_t - a.rast_ctl[0]);
        }
    }
    if (!mask) return;
    // Spike and event totals are not counted here: they follow exactly from
    // pn_total and the out-degrees on readback, so no wave contends for a
    // shared counter.
    if (LIST) {
        const uint32_t nsp = (uint32_t)__popcll(mask);
        uint32_t off = 0;
        if (lane == 0) off = atomicAdd(&a.cnt[L], nsp);
        off = uniform(off);
        if (spiked) {
            const uint32_t pos = off + wave_rank(mask);
            a.spk[(size_t)L * a.n_total + pos] = i;
            if (a.rast_ids) {
                const uint32_t Lp = (L + 2) % 3;
                const uint64_t base_t = a.totals[2 + Lp] + a.cnt[Lp];
                const uint64_t rp = base_t + pos - a.rast_ctl[0];
                if (rp < a.rast_cap_ids)
                    a.rast_ids[rp] = i;
                else
                    a.rast_ctl[2] = 1;
            }
        }
    }
    if (PUSH) {
        // push the rows of this wave's spikes, one spike at a time, all lanes
        uint64_t m = mask;
        while (m) {
            const uint32_t b = (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            const uint32_t lo = __builtin_amdgcn_readlane(row_lo, b);
            const uint32_t hi = __builtin_amdgcn_readlane(row_hi, b);
            push_row<RING, PACKED>(a, lo, hi, slot, lane);
            if (a.xs_lo) push_custom<RING>(a, wave * kWave + b, slot, lane);
        }
    }
}

// ---------------------------------------------------------------- scatter
// Two-kernel shape: one wave per spike of step t, all synapses of the row.
template <int RING, bool PACKED>
__global__ __launch_bounds__(kMaxBlock) void k_scatter(StepArgs a) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + uniform(threadIdx.x >> 6);
    const uint64_t t = *a.t_dev + a.k;
    const uint32_t L = (uint32_t)(t % 3);
    const uint32_t c = a.cnt[L];
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t q = wave; q < c; q += nwaves) {
        const uint32_t src = uniform(a.spk[(size_t)L * a.n_total + q]);
        const uint32_t lo = a.row_lo[src], hi = a.row_hi[src];
        push_row<RING, PACKED>(a, lo, hi, a.slot, lane);
        if (a.xs_lo) push_custom<RING>(a, src, a.slot, lane);
    }
}

// Watched-neuron state rows (sim/src/kernels.rs:141-147 moved out of the
// per-neuron loop): one thread per watched neuron after the step.
__global__ void k_watch(StepArgs a) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n_watch) return;
    const uint64_t t = *a.t_dev + a.k;
    a.xrec[(size_t)(t % a.xrec_depth) * a.n_watch + w] = a.x[a.watch[w]];
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

// Plain copy of n16 16-byte words (byte-counter calibration).
__global__ void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

void launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t st) {
    const uint64_t n16 = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n16);
}

// End of a chunk: advance the device step base.
__global__ void k_advance(uint64_t *t_dev, uint32_t k) { *t_dev += k; }

// ------------------------------------------------------------ dispatch
template <int RING, int PMODE, int MMODE, bool PUSH, bool PACKED, bool LIST>
static void go(const StepArgs &a, uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL((k_step<RING, PMODE, MMODE, PUSH, PACKED, LIST>), dim3(blocks),
                       dim3(a.block), 0, st, a);
}

template <int RING, int PMODE, int MMODE>
static void launch_m(const StepArgs &a, uint32_t blocks, bool push, bool packed, bool list,
                     hipStream_t st) {
    if (!push) {
        go<RING, PMODE, MMODE, false, false, true>(a, blocks, st);
    } else if (packed) {
        if (list) go<RING, PMODE, MMODE, true, true, true>(a, blocks, st);
        else go<RING, PMODE, MMODE, true, true, false>(a, blocks, st);
    } else {
        if (list) go<RING, PMODE, MMODE, true, false, true>(a, blocks, st);
        else go<RING, PMODE, MMODE, true, false, false>(a, blocks, st);
    }
}

template <int RING, int PMODE>
static void launch_p(const StepArgs &a, uint32_t blocks, int mmode, bool push, bool packed, bool list,
                     hipStream_t st) {
    switch (mmode) {
    case 0: launch_m<RING, PMODE, 0>(a, blocks, push, packed, list, st); break;
    case 1: launch_m<RING, PMODE, 1>(a, blocks, push, packed, list, st); break;
    default: launch_m<RING, PMODE, 2>(a, blocks, push, packed, list, st); break;
    }
}

template <int RING>
static void launch_r(const StepArgs &a, uint32_t blocks, int pmode, int mmode, bool push, bool packed,
                     bool list, hipStream_t st) {
    switch (pmode) {
    case 0: launch_p<RING, 0>(a, blocks, mmode, push, packed, list, st); break;
    case 1: launch_p<RING, 1>(a, blocks, mmode, push, packed, list, st); break;
    default: launch_p<RING, 2>(a, blocks, mmode, push, packed, list, st); break;
    }
}

void launch_step(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool push,
                 bool packed, bool list, hipStream_t st) {
    if (ring == 0) launch_r<0>(a, blocks, pmode, mmode, push, packed, list, st);
    else launch_r<1>(a, blocks, pmode, mmode, push, packed, list, st);
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
