"""Benchmark of the step hot path on one MI355X.

Default workload (N=1): the 100,000-neuron random-cube Izhikevich network
(~100 Poisson out-synapses each, ~10M synapses), dt 0.5 ms, synaptic tau
5 ms, Gaussian current sd 1 per step plus the reference's `--noise 1`
Poisson kicks (1 Hz x 20 mV), int-weights delay ring in Q16.16 (bit-exact
raster).  One "step" = one simulation step of the whole network; inputs
(network, state, ring) are resident in HBM before timing starts.

Prints ONE JSON line: synaptic events/s (whole job) with sim-ms/s, the
roofline of the dominant kernel (algorithmic bytes per launch over its
HIP-event-timed average launch duration) and the CPU baseline (the oracle,
one thread, on a bounded sample of the same workload).

    python bench.py [--steps 2000] [--warmup 20] [--settle 6000] [--n 100000]
                    [--ring int|float] [--no-cpu]
"""
import argparse
import gc
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, GB/s
B_NEURON = 33                   # algorithmic bytes per neuron-step
B_EVENT = 17                    # algorithmic bytes per synaptic event
PERSIST_MAX = 2048              # steps per persistent launch at most (library default)
METRIC = "synaptic events/s & sim-ms/s, 100k Izh; batched/served vs plain, same raster"


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def load_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (if any)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("TRAFFIC_JSON "):
            return json.loads(line[len("TRAFFIC_JSON "):])
    return None


def bench_batched(args, ring_int=True):
    """256 uneven networks (1k-20k neurons, half Izhikevich / half AdEx,
    own seeds, 8 Bernoulli(0.05) 1 ms bit-stream inputs each), 500 ms =
    1000 steps: one packed arena (launch shape, workgroup size and graph
    length picked by the autotuner) vs the same networks launched one at a
    time through the plain path with default settings.  Int ring: rasters
    compared per network (identical).  Float ring: the order of float ring
    sums varies between runs, so per-network spike counts are compared."""
    import numpy as np
    from neuro_edit_v2_amd import Sim, SimConfig
    from neuro_edit_v2_amd import workloads as W
    from neuro_edit_v2_amd.batch import Batch

    steps, warm, dt = 1000, 100, 0.5
    nets, series = W.batch_nets(args.batch_nets, dt=dt, steps=warm + steps)
    cfgs = [SimConfig(dt_ms=dt, synapse_tau_ms=5.0, seed=k + 1, int_weights=ring_int,
                      int_scale=65536.0, raster_steps=warm + steps) for k in range(len(nets))]
    b = Batch(nets, cfgs)
    for k in range(len(nets)):
        for q in range(8):
            b.play_input(k, q, series[k][:, q])
    tuned = None
    if not args.no_autotune:
        t_tune = time.perf_counter()
        rep = b.autotune(128)            # trial runs from step 0, then resets
        tuned = {"chosen": rep.strip().splitlines()[-1].replace("chosen: ", ""),
                 "tune_s": round(time.perf_counter() - t_tune, 2)}
    b.run_steps(warm)                    # untimed warm-up (graph capture, first touches)
    b.mark_counters()                    # window-start counters, taken on the device
    b.synchronize()
    t0 = time.perf_counter()
    b.run_steps(steps)
    b.synchronize()
    bwall = time.perf_counter() - t0
    bgpu = b.last_run_ms()
    bev = b.counts_since_mark()[1]
    plain_wall, plain_ev, same, spikes = 0.0, 0, True, 0
    dev = []
    for k, net in enumerate(nets if not args.batch_no_plain else []):
        s = Sim(net, cfgs[k])
        for q in range(8):
            s.play_input(q, series[k][:, q])
        s.run_steps(warm)
        s.mark_counters()
        s.synchronize()
        t1 = time.perf_counter()
        s.run_steps(steps)
        s.synchronize()
        plain_wall += time.perf_counter() - t1
        plain_ev += s.counts_since_mark()[1]
        _, pc, pi = s.read_spikes()
        _, bc, bi = b.read_spikes(k)
        same = same and np.array_equal(pc, bc) and np.array_equal(pi, bi)
        dev.append(abs(int(bc.sum()) - int(pc.sum())) / max(int(pc.sum()), 1))
        spikes += int(pc.sum())
        s.close()
    n_tot = sum(int(n["n"]) for n in nets)
    out = {"workload": f"{len(nets)} grid liquids, {n_tot} neurons, "
                        f"{sum(len(n['tgt']) for n in nets)} synapses, half izh / half adex, "
                        f"8 Bernoulli(0.05) 1 ms inputs each, {steps} steps (500 ms) timed after "
                        f"{warm} warm-up steps, {'int' if ring_int else 'float'} ring",
            "autotune": tuned,
            "batched_us_per_step": round(bgpu / steps * 1e3, 3),
            "batched_events_per_s": round(bev / bwall, 1),
            "plain_sequential_events_per_s": round(plain_ev / plain_wall, 1) if plain_wall else None,
            "speedup": round((bev / bwall) / (plain_ev / plain_wall), 2) if plain_wall else None,
            "batched_wall_s": round(bwall, 5), "batched_gpu_ms": round(bgpu, 3),
            "plain_wall_s": round(plain_wall, 5),
            "batched_sim_ms_per_s": round(len(nets) * steps * dt / bwall, 1),
            "events": int(bev), "events_plain": int(plain_ev), "spikes": spikes,
            # whole-step roofline: the algorithmic bytes of a step (33 B per
            # neuron, 17 B per event, as the headline) over the step's GPU time
            "roofline_frac": round((n_tot * B_NEURON + bev / steps * B_EVENT)
                                   / (bgpu / steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if args.batch_no_plain:
        pass
    elif ring_int:
        out["rasters_identical"] = bool(same and bev == plain_ev)
    else:
        # float sums in atomic order: per-network spike totals, batched vs plain
        out["spike_count_rel_dev_mean"] = round(float(np.mean(dev)), 5)
        out["spike_count_rel_dev_max"] = round(float(np.max(dev)), 5)
    return out


def bench_served(args, ring_int=True):
    """LSM trial requests (grid liquids of edge U{10..37}, 8 bit streams, 10
    windows x 50 ms, one of 4 stream pairs sharing bits per window, 500 ms
    each) arriving as a Poisson process at 50 req/s; continuous batching in
    one resident arena.  Latency = completion - arrival; a few requests are
    re-run alone through the plain path and their window counts compared."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from neuro_edit_v2_amd import Sim, SimConfig
    from neuro_edit_v2_amd import workloads as W
    from neuro_edit_v2_amd.serve import Server, window_counts

    n_req, rate, dt, steps, win = args.serve_requests, 50.0, 0.5, 1000, 100
    with ThreadPoolExecutor(8) as ex:
        reqs = list(ex.map(lambda k: W.lsm_request(5000 + k, dt=dt), range(n_req)))
    cfgs = [SimConfig(dt_ms=dt, synapse_tau_ms=5.0, seed=5000 + k, int_weights=ring_int,
                      int_scale=65536.0, raster_steps=steps) for k in range(n_req)]
    srv = Server(max_requests=256, max_delay=15, n_windows=10, chunk_steps=96,
                 int_weights=int(ring_int), int_scale=65536.0)
    arr = np.cumsum(np.random.default_rng(77).exponential(1.0 / rate, n_req))
    ids = []
    t_start = time.perf_counter()
    for k, (net, ser, _) in enumerate(reqs):
        delay = t_start + arr[k] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(net, cfgs[k], ser, steps, win))
    got = [srv.wait(r, timeout_ms=120000) for r in ids]
    st = srv.stats()
    lat = np.array([(g[1]["done_ns"] - g[1]["submit_ns"]) * 1e-6 for g in got])
    span = (max(g[1]["done_ns"] for g in got) - min(g[1]["submit_ns"] for g in got)) * 1e-9
    # alone == in the mix: a sample of requests re-run alone through the
    # plain path (int: window counts identical; float: the float ring sums
    # in atomic order, so the per-neuron window counts are compared by
    # their mean absolute deviation)
    same, mads = True, []
    n_chk = min(args.serve_check, n_req)
    for k in range(n_chk):
        net, ser, _ = reqs[k]
        s = Sim(net, cfgs[k])
        for q in range(8):
            s.play_input(q, ser[:, q])
        s.run_steps(steps)
        _, c, i = s.read_spikes()
        alone = window_counts(c, i, int(net["n"]), win, 10)
        if ring_int:
            same = same and np.array_equal(alone, got[k][0])
        else:
            mads.append(float(np.abs(alone.astype(np.int64) - got[k][0]).mean()
                              / max(alone.mean(), 1e-9)))
        s.close()
    srv.close()
    return {"requests": n_req, "arrival_rate_hz": rate,
            "neurons_mean": round(float(np.mean([r[0]["n"] for r in reqs])), 1),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "events_per_s": round(st["events"] / span, 1), "events": int(st["events"]),
            "span_s": round(span, 3), "peak_in_flight": int(st["peak_active"]),
            "gpu_busy_frac": round(st["gpu_ms"] / 1e3 / span, 4),
            "ring": "int Q16.16" if ring_int else "float",
            "alone_vs_mix_checked": n_chk,
            **({"alone_equals_mix": bool(same)} if ring_int else
               {"alone_vs_mix_count_mad_mean": round(float(np.mean(mads)), 5) if mads else None,
                "alone_vs_mix_count_mad_max": round(float(np.max(mads)), 5) if mads else None})}


REF_PUBLISHED = {"hardware": "RTX 3090 (reference CUDA backend)", "wall_s": 0.549,
                 "realtime_x": 5.5, "source": "NOTES-for-data.md:24-30"}


def bench_reference_regime(args):
    """The reference's own benchmark regime (sim/src/bench.rs:10-89,
    NOTES-for-data.md:24-30): bench_net(100000, K=100, seed 1, wscale 0.7,
    w-inh 5), dt 0.1 ms, synaptic tau 2 ms, Poisson kicks 500 Hz x 20 mV
    (the noise of the published 10k runs, NOTES-for-data.md:2391; the 21 Hz
    regime), 100 ms warm-up then 3000 ms timed, as run_bench does.  Float
    ring (the reference default) and int ring (scale 4096).  The published
    figure for this size (0.549 s wall, 5.5x realtime on an RTX 3090) is
    context from other hardware, not a target.  Roofline as the headline:
    33 B per neuron-step and 17 B per event over the step's GPU time --
    about 210 spikes and 21k deliveries per step, so the scatter is loaded."""
    from neuro_edit_v2_amd import Sim, SimConfig
    from neuro_edit_v2_amd import net as N

    n, dt, warm_ms, sim_ms = args.ref_n, 0.1, 100.0, 3000.0
    t_gen = time.time()
    net = N.bench_net(n, 100, 1, 0.7, 5.0, dt)
    t_gen = time.time() - t_gen
    out = {"workload": f"bench_net({n}, K=100, seed 1, wscale 0.7, w-inh 5), dt 0.1 ms, tau 2 ms, "
                       f"kicks 500 Hz x 20 mV, {warm_ms:g} ms warm-up then {sim_ms:g} ms timed",
           "neurons": n, "synapses": int(len(net["tgt"])), "gen_s": round(t_gen, 2),
           "published": REF_PUBLISHED}
    steps = int(round(sim_ms / dt))
    for ring in ("float", "int"):
        sim = Sim(net, SimConfig(dt_ms=dt, synapse_tau_ms=2.0, noise_rate_hz=500.0,
                                 noise_kick_mv=20.0, seed=1, int_weights=ring == "int",
                                 int_scale=4096.0))
        sim.run_ms(warm_ms)
        sim.mark_counters()
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run_steps(steps)
        sim.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = sim.last_run_ms()
        spikes, events = sim.counts_since_mark()
        step_s = gpu_ms * 1e-3 / steps
        frac = (n * B_NEURON + events / steps * B_EVENT) / step_s / 1e9 / HBM_PEAK_GBS
        out[ring] = {"wall_s": round(wall, 4), "realtime_x": round(sim_ms / 1000.0 / wall, 2),
                     "us_per_step": round(wall / steps * 1e6, 3),
                     "gpu_us_per_step": round(step_s * 1e6, 3),
                     "events_per_s": round(events / wall, 1),
                     "sim_ms_per_s": round(sim_ms / wall, 1),
                     "spikes_timed": int(spikes), "spikes_lifetime": int(sim.spikes_total()),
                     "rate_hz": round(spikes / n / (sim_ms / 1000.0), 2),
                     "events_per_step": round(events / steps, 1),
                     "step_shape": ["two-kernel", "fused", "paired", "persistent"][sim.launch_shape()],
                     "roofline_frac": round(frac, 4)}
        sim.close()
    return out


def bench_large(args, kw):
    """The 1M-neuron network (random cube, Poisson(100) out-degree, ~100M
    synapses), same generator, noise and ring as the headline workload,
    400 steps after 2000 warm-up steps: the step kernel at a size where the
    chip is full, with its roofline fraction."""
    from neuro_edit_v2_amd import Sim, SimConfig
    from neuro_edit_v2_amd import net as N

    n, steps, warm, dt = args.large_n, 400, 2000, kw["dt_ms"]
    t_gen = time.time()
    net = N.cube_net(n, mean_deg=100.0, seed=1, w_scale=0.7)
    t_gen = time.time() - t_gen
    n_syn = int(len(net["tgt"]))
    sim = Sim(net, SimConfig(exec_mode=args.exec_mode, step_shape=args.shape, **kw))
    del net
    sim.run_steps(warm)
    sim.mark_counters()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run_steps(steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = sim.last_run_ms()
    tail_ms, tail_n = sim.last_run_tail()
    spikes, events = sim.counts_since_mark()
    ev_step = events / steps
    spl = 2 if sim.launch_shape() == 2 else 1     # steps per launch
    kbytes = spl * (n * B_NEURON + ev_step * B_EVENT)
    kavg_ms = tail_ms / tail_n if tail_n else gpu_ms / (steps / spl)
    achieved = kbytes / (kavg_ms * 1e-3) / 1e9
    traffic = load_traffic(f"traffic_fused_cube{n // 1000}k_{'int' if kw['int_weights'] else 'float'}.txt")
    return {"workload": f"random-cube Izhikevich {n} neurons x Poisson(100) out-synapses, "
                        f"dt 0.5 ms, {steps} steps, same noise and ring",
            "neurons": n, "synapses": n_syn, "gen_s": round(t_gen, 2),
            "dev_mem_bytes": int(sim.dev_mem_bytes()),
            "steps": steps, "warmup": warm,
            "events_per_s": round(events / wall, 1),
            "sim_ms_per_s": round(steps * dt / wall, 1),
            "ms_per_step": round(wall / steps * 1e3, 5),
            "spikes": int(spikes), "events": int(events),
            "rate_hz": round(spikes / n / (steps * dt / 1000.0), 3),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": traffic.get("bytes_per_launch") if traffic else None,
                         "steps_per_launch": spl, "avg_launch_us": round(kavg_ms * 1e3, 3),
                         "algorithmic_bytes_per_launch": int(kbytes)}}


def settle(sim, steps, total):
    """Clock-settle phase, untimed: the timed call run_steps(steps) repeated
    until at least `total` steps have run.  Over the first ~6000 back-to-back
    steps the device ramps to its steady clocks (100k: 2.85 -> 2.76 us per
    step, flat after), so a short --warmup alone would time the
    ramp; and the library replays a chunk from a captured graph once it has
    seen that chunk's length before, which the repeats give the timed call.
    The last repeats each end in a synchronisation, as the timed call does,
    so that the runtime's completion bookkeeping of the thousands of
    launches before them is not left to the timed call (a single 20-step
    call after the unsynchronised repeats measured 237 us of wall time
    against 86 us typical).  Returns the steps run."""
    reps = max(2, -(-total // max(steps, 1)))
    for _ in range(reps):
        sim.run_steps(steps)
    for _ in range(5):
        sim.run_steps(steps)
        sim.synchronize()
    return (reps + 5) * steps


def host_spin_wait():
    """Ask the HIP runtime to spin-wait for completion (hipDeviceScheduleSpin)
    in this process.  Set before the first HIP call (a flag of the device's
    context): a synchronisation then returns when the device finishes rather
    than when the host thread is woken from a blocking wait, so a short timed
    call's wall time does not carry the host's wake-up jitter.  The device
    work is the same.  Returns the runtime's status (0: set)."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    return int(hip.hipSetDeviceFlags(ctypes.c_uint(1)))


def main():
    spin = host_spin_wait()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--ring", choices=["int", "float"], default="int")
    ap.add_argument("--shape", type=int, default=-1, help="-1 auto, 0 two kernels, 1 fused")
    ap.add_argument("--exec", dest="exec_mode", type=int, default=0)
    ap.add_argument("--profile-steps", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=8000,
                    help="oracle steps timed for the CPU baseline (about 10 s of CPU work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many bytes (byte-counter calibration)")
    ap.add_argument("--lane-bytes", type=int, default=16, choices=[4, 16])
    ap.add_argument("--launch-probe", type=int, default=0,
                    help="only time graph replays of empty kernels of this many blocks")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--batch-nets", type=int, default=256)
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched form: keep the default launch shape")
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="batched form: skip the one-at-a-time runs (for profiling the packed kernel)")
    ap.add_argument("--skip-served", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--skip-large", action="store_true")
    ap.add_argument("--skip-other-ring", action="store_true",
                    help="skip the run of the same network with the other ring")
    ap.add_argument("--large-n", type=int, default=1000000)
    ap.add_argument("--skip-ref", action="store_true",
                    help="skip the reference's own benchmark regime (bench_net, dt 0.1)")
    ap.add_argument("--ref-n", type=int, default=100000)
    ap.add_argument("--serve-check", type=int, default=10,
                    help="served requests re-run alone and compared with their served result")
    ap.add_argument("--settle", type=int, default=6000,
                    help="untimed clock-settle steps after the warm-up, run as repeats of the "
                         "timed call (the device reaches its steady clocks over about 6000 "
                         "back-to-back steps of this network)")
    args = ap.parse_args()
    if args.copy_probe or args.launch_probe:
        import ctypes
        from neuro_edit_v2_amd import _lib
        v = ctypes.c_double()
        if args.copy_probe:
            _lib.check(_lib.lib().nh_probe_copy(args.copy_probe, 20, args.lane_bytes, ctypes.byref(v)))
            print(json.dumps({"copy_bytes": args.copy_probe, "lane_bytes": args.lane_bytes,
                              "avg_ms": v.value,
                              "gb_per_s": 2 * args.copy_probe / (v.value * 1e-3) / 1e9}))
        if args.launch_probe:
            for blocks in (1, args.launch_probe):
                _lib.check(_lib.lib().nh_probe_launch(blocks, 2000, ctypes.byref(v)))
                print(json.dumps({"empty_kernel_graph_us": v.value, "blocks": blocks,
                                  "threads_per_block": 256}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this backend runs on one GPU; multi-GPU figures are not produced "
              "(the networks of a batch are independent units and would shard by network)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from neuro_edit_v2_amd import Sim, SimConfig
    from neuro_edit_v2_amd import net as N

    t_gen = time.time()
    net = N.cube_net(args.n, mean_deg=100.0, seed=1, w_scale=0.7)
    t_gen = time.time() - t_gen
    deg = np.diff(net["src_ptr"].astype(np.int64))
    dt = 0.5
    kw = dict(dt_ms=dt, synapse_tau_ms=5.0, noise_sd=1.0, noise_rate_hz=1.0, noise_kick_mv=20.0,
              seed=1, int_weights=args.ring == "int", int_scale=65536.0)
    t_build = time.perf_counter()
    sim = Sim(net, SimConfig(exec_mode=args.exec_mode, step_shape=args.shape, **kw))
    sim.synchronize()
    t_build = time.perf_counter() - t_build      # host CSR -> HBM upload and set-up
    # The window-start counters are a device-side snapshot enqueued behind
    # the warm-up: reading them on the host here would leave the device idle
    # for a host round trip right before the timed steps (at 1M neurons a
    # 1.2 ms read made the next 400 steps 4-5 % slower: the clocks drop).
    # Warm-up (--warmup steps), then a clock-settle phase of its own
    # (see settle()).
    sim.run_steps(args.warmup)
    settled = settle(sim, args.steps, args.settle)
    sim.mark_counters()
    sim.synchronize()
    gc.disable()                         # no collector pause inside the timed region
    t0 = time.perf_counter()
    sim.run_steps(args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    gc.enable()
    gpu_ms = sim.last_run_ms()
    tail_ms, tail_n = sim.last_run_tail()
    spikes, events = sim.counts_since_mark()
    value = events / wall

    # Dominant-kernel roofline.  Fused shape: the step kernel is the only
    # per-step launch, so its average launch duration is the timed region's
    # HIP-event time (events on the sim's stream) over its launches, taken
    # from the end of the region's first launch to the end of the region
    # (last_run_tail): the launches back to back, without the host's latency
    # before the first kernel; it includes every inter-launch gap and the
    # chunk's closing step-counter advance, so it does not flatter the
    # kernel.  Two-kernel
    # shape: a profiled pass with an event pair around every launch splits
    # the step between the two kernels.
    ev_per_step = events / args.steps
    shape = sim.launch_shape()           # 0 two kernels, 1 fused, 2 paired
    fused = shape != 0
    prof = None
    if args.profile_steps:
        ev1 = sim.events_total()
        main_ms, main_n, aux_ms, aux_n = sim.profile_steps(args.profile_steps)
        prof = {"main_us": round(main_ms / max(main_n, 1) * 1e3, 3),
                "aux_us": round(aux_ms / max(aux_n, 1) * 1e3, 3) if aux_n else None,
                "note": "direct launches, event pair per launch (adds dispatch latency)"}
        fused = aux_n == 0
    # steps per launch of the dominant kernel: the paired kernel runs two
    # (every launch of the timed region is a pair: chunks and the timed
    # region hold even step counts); the persistent kernel runs the timed
    # region in launches of up to PERSIST_MAX steps
    spl = 2 if shape == 2 else (min(args.steps, PERSIST_MAX) if shape == 3 else 1)
    if fused:
        kname = {2: "k_pair (two steps per launch: update + own-spike push, delay-1 arrivals pulled)",
                 3: "k_run (persistent: all steps of a run in one launch, delay-1 arrivals pulled)"}.get(
                     shape, "k_step (fused: update + own-spike push)")
        kbytes = spl * (args.n * B_NEURON + ev_per_step * B_EVENT)
        kavg = tail_ms / tail_n if tail_n else gpu_ms / (args.steps / spl)
    elif main_ms >= aux_ms:
        kname = "k_step (update + compaction)"
        kbytes = args.n * B_NEURON
        kavg = main_ms / max(main_n, 1)
    else:
        kname = "k_scatter"
        kbytes = ev_per_step * B_EVENT
        kavg = aux_ms / max(aux_n, 1)
    achieved = kbytes / (kavg * 1e-3) / 1e9
    tname = {2: "paired", 3: "persistent"}.get(shape, "fused" if fused else "neuron")
    traffic = load_traffic(f"traffic_{tname}_cube{args.n // 1000}k_{args.ring}.txt")
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "kernel": kname, "steps_per_launch": spl, "avg_launch_us": round(kavg * 1e3, 3),
            "timed_launches": tail_n if (fused and tail_n) else None,
            "avg_launch_window": ("timed region, end of its first step launch to its end (launches back "
                                  "to back)" if (fused and tail_n) else "timed region, whole"),
            "algorithmic_bytes_per_launch": int(kbytes),
            "bytes_per_neuron_step": B_NEURON, "bytes_per_event": B_EVENT,
            "events_per_step": round(ev_per_step, 1),
            "whole_step_frac": round((args.n * B_NEURON + ev_per_step * B_EVENT)
                                     / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
            "per_launch_profile": prof}

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ref = O.OracleSim(net, O.default_cfg(**kw))
        warm = min(200, args.cpu_steps)
        ref.run_steps(warm)
        cev0 = ref.events_total()
        c0 = time.perf_counter()
        ref.run_steps(args.cpu_steps)
        cwall = time.perf_counter() - c0
        cev = int(ref.events_total() - cev0)
        cpu = {"value": round(cev / cwall, 1), "unit": "synaptic events/s", "cores": 1,
               "kind": "port",
               "sample": f"same {args.n}-neuron network and config, {args.cpu_steps} steps after "
                         f"{warm} warm-up steps, oracle C restatement of CpuSim, 1 thread "
                         f"({cpu_model()})",
               "sim_ms_per_s": round(args.cpu_steps * dt / cwall, 2),
               "wall_s": round(cwall, 3)}

    del sim
    # the same network with the other ring (float if the headline ran int):
    # the float-ring run of the 100k configuration beside the int-weights one
    other = None
    if not args.skip_other_ring:
        kw2 = dict(kw, int_weights=not kw["int_weights"])
        sim2 = Sim(net, SimConfig(exec_mode=args.exec_mode, step_shape=args.shape, **kw2))
        sim2.run_steps(args.warmup)
        settle(sim2, args.steps, args.settle)
        sim2.mark_counters()
        sim2.synchronize()
        t1 = time.perf_counter()
        sim2.run_steps(args.steps)
        sim2.synchronize()
        w2 = time.perf_counter() - t1
        ev2 = sim2.counts_since_mark()[1]
        sh2 = sim2.launch_shape()
        spl2 = 2 if sh2 == 2 else (min(args.steps, PERSIST_MAX) if sh2 == 3 else 1)
        kb2 = spl2 * (args.n * B_NEURON + ev2 / args.steps * B_EVENT)
        t2_ms, t2_n = sim2.last_run_tail()
        ka2 = t2_ms / t2_n if t2_n else sim2.last_run_ms() / (args.steps / spl2)
        other = {"ring": "int Q16.16" if kw2["int_weights"] else "float",
                 "value": round(ev2 / w2, 1), "unit": "synaptic events/s",
                 "ms_per_step": round(w2 / args.steps * 1e3, 5),
                 "sim_ms_per_s": round(args.steps * dt / w2, 1),
                 "events": int(ev2),
                 "roofline_frac": round(kb2 / (ka2 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "avg_launch_us": round(ka2 * 1e3, 3), "steps_per_launch": spl2}
        del sim2
    large = None if args.skip_large else bench_large(args, kw)
    ref_regime = None if args.skip_ref else bench_reference_regime(args)
    batched = None
    if not args.skip_batched:
        batched = {"int": bench_batched(args, True), "float": bench_batched(args, False)}
    served = None
    if not args.skip_served:
        served = {"int": bench_served(args, True), "float": bench_served(args, False)}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, i32 ring" if args.ring == "int" else "f32",
        "data": "synthetic",
        "config": {"workload": f"random-cube Izhikevich {args.n} neurons x Poisson(100) "
                               f"out-synapses, dt 0.5 ms, 1000 ms per 2000 steps, "
                               f"{args.ring}-weights ring"
                               + (" Q16.16" if args.ring == "int" else ""),
                   "neurons": args.n, "synapses": int(len(net["tgt"])),
                   "noise": "gaussian sd 1/step + poisson kicks 1 Hz x 20 mV",
                   "step_shape": ["two-kernel", "fused", "paired", "persistent"][shape],
                   "gen_s": round(t_gen, 2)},
        "sim_ms_per_s": round(args.steps * dt / wall, 1),
        "spikes": int(spikes),
        "events": int(events),
        "rate_hz": round(spikes / args.n / (args.steps * dt / 1000.0), 3),
        "gpu_event_ms": round(gpu_ms, 3),
        # the timed steps' GPU time (HIP events on the sim's stream) per
        # step, beside the wall-clock figure above (which also holds the
        # host's launch and synchronisation of the timed region)
        "gpu_ms_per_step": round(gpu_ms / args.steps, 6),
        "clock_settle_steps": settled,
        "host_sync": "spin" if spin == 0 else "runtime default",
        # the network arrives as host arrays: the rate with its upload counted
        "build_s": round(t_build, 4),
        "events_per_s_incl_build": round(events / (wall + t_build), 1),
        "roofline": roof,
        "cpu_baseline": cpu,
        "other_ring": other,
        "large": large,
        "reference_regime": ref_regime,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
