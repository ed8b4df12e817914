// kernels.hip -- the per-step hot path for gfx950 (wave64, CDNA4).
//
// Two execution shapes, both parity-equal to the CPU step order of
// sim/src/cpu.rs:192-306:
//
//  * two-phase: k_neuron (ring read/clear, decay, noise, model step,
//    wave64 ballot compaction into the step's spike list) then k_scatter
//    (one wave per spike, lanes over the delay-sorted CSR row, hardware
//    atomics into the delay ring);
//
//  * fused: ONE kernel per step.  Synapses with delay >= 2 of the previous
//    step's spikes are pushed (they land in slots this step never reads),
//    delay-1 contributions are pulled per target from a target-major list
//    and the previous step's spike bitmap.  No grid-wide barrier is needed
//    inside a step.
//
// Spike lists rotate over three buffers (step t uses list t%3); counter and
// cumulative-total bookkeeping is done by the first thread of the grid
// without races (see k_neuron).
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

// Push the out-synapses of one spike into the ring: lanes stride the row
// (delay-sorted, so a contiguous run of lanes hits one ring row).
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

// ---------------------------------------------------------------- neuron
template <int RING, int PMODE, int MMODE, bool FUSED, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_neuron(StepArgs a) {
    using RT = typename Ring<RING>::T;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & (kWave - 1);
    const uint32_t wib = uniform(tid >> 6);
    const uint32_t wave = blockIdx.x * (kBlock / kWave) + wib;
    const uint32_t i = wave * kWave + lane;
    const uint64_t t = *a.t_dev + a.k;
    const uint32_t D = a.D;
    const uint32_t slot = (uint32_t)(t % D);
    const uint32_t L = (uint32_t)(t % 3), Lp = (L + 2) % 3, Ln = (L + 1) % 3;
    // Cumulative spikes before step t = before step t-1 + count of t-1;
    // both are final (written by earlier launches), so every wave may read
    // them.  fill[] lives in totals[2..4].
    const uint64_t base_t = a.totals[2 + Lp] + a.cnt[Lp];
    if (blockIdx.x == 0 && tid == 0) {
        a.totals[2 + L] = base_t;
        a.cnt[Ln] = 0;
        if (a.rast_off) {
            const uint64_t rs = t - a.rast_ctl[1];
            if (rs < a.rast_cap_steps) a.rast_off[rs] = (uint32_t)(base_t - a.rast_ctl[0]);
        }
    }

    // Fused shape: push the previous step's delay>=2 synapses first.
    if (FUSED && t > 0) {
        const uint32_t cprev = a.cnt[Lp];
        const uint32_t nwaves = gridDim.x * (kBlock / kWave);
        const uint32_t slotp = (slot + D - 1) % D;  // (t-1) % D
        uint64_t ev = 0;
        for (uint32_t q = wave; q < cprev; q += nwaves) {
            const uint32_t src = uniform(a.spk[(size_t)Lp * a.n_total + q]);
            const uint32_t lo = a.src_ptr[src], hi = a.src_ptr[src + 1];
            push_row<RING, PACKED>(a, lo + a.skip1[src], hi, slotp, lane);
            ev += hi - lo;
        }
        if (lane == 0 && ev) atomicAdd((unsigned long long *)&a.totals[1], (unsigned long long)ev);
    }

    // waves past the arena (last block) only helped with the push above
    if (wave >= a.words) return;
    const uint32_t net = a.single_net ? 0u : uniform((uint32_t)a.wave_net[wave]);
    const NetRec &nr = a.nets[net];
    const uint32_t local = i - nr.base;
    const bool act = local < nr.n && nr.active;
    bool spiked = false;
    if (act) {
        RT *ring = (RT *)a.ring;
        const size_t ri = (size_t)slot * a.n_total + i;
        RT rv = ring[ri];
        ring[ri] = (RT)0;
        if (FUSED) {
            // delay-1 arrivals from the previous step, pulled
            const uint64_t *pb = a.pbits + (size_t)((t + 1) & 1) * a.words;
            const RT *w1 = (const RT *)a.in1_w;
            for (uint32_t j = a.in1_ptr[i]; j < a.in1_ptr[i + 1]; ++j) {
                const uint32_t src = a.in1_src[j];
                if ((pb[src >> 6] >> (src & 63)) & 1ull) rv += w1[j];
            }
        }
        const float arr = Ring<RING>::get(rv, nr.int_scale);
        const float sv = a.s[i] + arr;
        float I = sv + (a.ext ? a.ext[i] : 0.0f);
        if (a.stim_ptr) {
            const float *row = a.stim_win + (size_t)a.k * a.stim_cols + nr.stim_base;
            for (uint32_t j = a.stim_ptr[i]; j < a.stim_ptr[i + 1]; ++j)
                I += a.stim_w[j] * row[a.stim_input[j]];
        }
        float pr[kMaxParams];
        uint32_t model;
        if (PMODE == 0) {
            const PRow &r = a.ptab[a.prow[i]];
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
        const float gain = is_adex ? pr[0] : 1.0f;
        const uint64_t tl = t - (uint64_t)nr.t0;
        if (nr.pkick > 0.0f && kick_fires(local, tl, nr.seed, nr.pkick)) I += nr.kick * gain;
        if (nr.sd != 0.0f) I += (nr.sd * gauss_draw(local, tl, nr.seed)) * gain;
        float2 xs = a.x[i];
        const float dt = nr.dt;
        if (is_adex)
            spiked = adex_step(pr, xs, I, dt);
        else
            spiked = izh_step(pr[0], pr[1], pr[2], pr[3], xs, I, dt);
        a.x[i] = xs;
        a.s[i] = sv * nr.decay_s;
        if (spiked) {
            a.pn_total[i] += 1;
            a.last_spike[i] = t < 0xFFFFFFFEull ? (uint32_t)t : 0xFFFFFFFEu;
        }
    }

    // wave64 ballot: the mask IS this wave's word of the spike bitmap
    const uint64_t mask = __ballot(spiked);
    if (lane == 0) {
        if (a.bits) a.bits[(size_t)(t % a.bits_rows) * a.words + wave] = mask;
        if (a.pbits) a.pbits[(size_t)(t & 1) * a.words + wave] = mask;
    }
    const uint32_t nsp = (uint32_t)__popcll(mask);
    if (nsp) {
        uint32_t off = 0;
        if (lane == 0) off = atomicAdd(&a.cnt[L], nsp);
        off = uniform(off);
        if (spiked) {
            const uint32_t pos = off + wave_rank(mask);
            a.spk[(size_t)L * a.n_total + pos] = i;
            if (a.rast_ids) {
                const uint64_t rp = base_t + pos - a.rast_ctl[0];
                if (rp < a.rast_cap_ids)
                    a.rast_ids[rp] = i;
                else
                    a.rast_ctl[2] = 1;
            }
        }
        if (lane == 0 && a.net_spikes)
            atomicAdd((unsigned long long *)&a.net_spikes[net], (unsigned long long)nsp);
    }
}

// ---------------------------------------------------------------- scatter
// Two-phase shape: one wave per spike of step t, all synapses of the row.
template <int RING, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_scatter(StepArgs a) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = blockIdx.x * (kBlock / kWave) + uniform(threadIdx.x >> 6);
    const uint64_t t = *a.t_dev + a.k;
    const uint32_t L = (uint32_t)(t % 3);
    const uint32_t slot0 = (uint32_t)(t % a.D);
    const uint32_t c = a.cnt[L];
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    uint64_t ev = 0;
    for (uint32_t q = wave; q < c; q += nwaves) {
        const uint32_t src = uniform(a.spk[(size_t)L * a.n_total + q]);
        const uint32_t lo = a.src_ptr[src], hi = a.src_ptr[src + 1];
        push_row<RING, PACKED>(a, lo, hi, slot0, lane);
        ev += hi - lo;
    }
    if (lane == 0 && ev) atomicAdd((unsigned long long *)&a.totals[1], (unsigned long long)ev);
}

// Watched-neuron state rows (sim/src/kernels.rs:141-147 moved out of the
// per-neuron loop): one thread per watched neuron after the step.
__global__ void k_watch(StepArgs a) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n_watch) return;
    const uint64_t t = *a.t_dev + a.k;
    a.xrec[(size_t)(t % a.xrec_depth) * a.n_watch + w] = a.x[a.watch[w]];
}

// End of a chunk: advance the device step base.
__global__ void k_advance(uint64_t *t_dev, uint32_t k) { *t_dev += k; }

// ------------------------------------------------------------ dispatch
using KFn = void (*)(StepArgs);

template <int RING, int PMODE, int MMODE, bool FUSED, bool PACKED>
static void launch_neuron_t(const StepArgs &a, uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL((k_neuron<RING, PMODE, MMODE, FUSED, PACKED>), dim3(blocks), dim3(kBlock), 0,
                       st, a);
}

template <int RING, int PMODE, int MMODE>
static void launch_neuron_f(const StepArgs &a, uint32_t blocks, bool fused, bool packed,
                            hipStream_t st) {
    if (fused) {
        if (packed) launch_neuron_t<RING, PMODE, MMODE, true, true>(a, blocks, st);
        else launch_neuron_t<RING, PMODE, MMODE, true, false>(a, blocks, st);
    } else {
        launch_neuron_t<RING, PMODE, MMODE, false, false>(a, blocks, st);
    }
}

template <int RING, int PMODE>
static void launch_neuron_m(const StepArgs &a, uint32_t blocks, int mmode, bool fused, bool packed,
                            hipStream_t st) {
    switch (mmode) {
    case 0: launch_neuron_f<RING, PMODE, 0>(a, blocks, fused, packed, st); break;
    case 1: launch_neuron_f<RING, PMODE, 1>(a, blocks, fused, packed, st); break;
    default: launch_neuron_f<RING, PMODE, 2>(a, blocks, fused, packed, st); break;
    }
}

void launch_neuron(const StepArgs &a, uint32_t blocks, int ring, int pmode, int mmode, bool fused,
                   bool packed, hipStream_t st) {
    if (ring == 0) {
        if (pmode == 0) launch_neuron_m<0, 0>(a, blocks, mmode, fused, packed, st);
        else launch_neuron_m<0, 1>(a, blocks, mmode, fused, packed, st);
    } else {
        if (pmode == 0) launch_neuron_m<1, 0>(a, blocks, mmode, fused, packed, st);
        else launch_neuron_m<1, 1>(a, blocks, mmode, fused, packed, st);
    }
}

void launch_scatter(const StepArgs &a, uint32_t blocks, int ring, bool packed, hipStream_t st) {
    if (ring == 0) {
        if (packed) hipLaunchKernelGGL((k_scatter<0, true>), dim3(blocks), dim3(kBlock), 0, st, a);
        else hipLaunchKernelGGL((k_scatter<0, false>), dim3(blocks), dim3(kBlock), 0, st, a);
    } else {
        if (packed) hipLaunchKernelGGL((k_scatter<1, true>), dim3(blocks), dim3(kBlock), 0, st, a);
        else hipLaunchKernelGGL((k_scatter<1, false>), dim3(blocks), dim3(kBlock), 0, st, a);
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
