"""Two steps per launch (step_shape=2): the delay-1 arrivals of the first
step of a launch are pulled by their targets' waves instead of pushed.

Int ring: raster, v, u and I are the oracle's bit for bit whatever the mix
of delays -- few delay-1 synapses (random cube), many (grid liquid with
local wiring, stimulus inputs, Izhikevich and AdEx), all of them, with more
than 64 delay-1 synapses into one wave -- over odd step counts (a one-step
launch closes each chunk), with graph and direct launches.  Float ring:
within the tolerances of the fused shape.  Where the shape cannot run
(watched neurons) the step falls back to one launch per step.  The bounded
wait's failure path, forced from the host, ends the launch and reports.
"""
import ctypes as C

import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import _lib as L
from neuro_edit_v2_amd import net as N
from neuro_edit_v2_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG0 = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_sd=1.0, noise_rate_hz=20.0, noise_kick_mv=20.0,
            seed=7)


def all_delay1_net(n, deg, seed):
    """Random targets, every synapse of delay 1 (the reference's default
    delay at dt = delay_ms), signed weights."""
    r = np.random.default_rng(seed)
    E = n * deg
    w = r.uniform(0.2, 2.5, E).astype(np.float32)
    w[r.random(E) < 0.2] *= -1.0
    return dict(n=n, src_ptr=np.arange(0, E + 1, deg, dtype=np.uint32),
                tgt=r.integers(0, n, E, dtype=np.uint32), w=w,
                delay=np.ones(E, np.uint16),
                model_ids=np.zeros(1, np.uint8), model=np.zeros(n, np.uint8), n_params=4,
                n_state=2, p=np.tile(np.float32([0.02, 0.2, -65.0, 8.0]), n),
                x0=np.tile(np.float32([-65.0, -13.0]), n))


def delay1_per_wave(net):
    d1 = net["delay"] == 1
    return np.bincount(net["tgt"][d1] // 64, minlength=(net["n"] + 63) // 64)


def grid_case(model):
    net = N.grid_net(0, model=model, seed=3 + model, dt_ms=0.5, dims=(17, 16, 15))
    bits = W.stream_window(W.rng_stream(50 + model, 8, 0.05), 8, 301)
    return net, bits


def check_int(net, steps, kw, series=None, exec_mode=2, exp_mode=0):
    sim = Sim(net, SimConfig(raster_steps=steps, step_shape=2, exec_mode=exec_mode, **kw))
    assert sim.launch_shape() == 2
    ref = O.OracleSim(net, O.default_cfg(exp_mode=exp_mode, **kw))
    if series is not None:
        for q in range(series.shape[1]):
            sim.play_input(q, series[:, q])
            ref.play_input(q, series[:, q])
    sim.run_steps(steps)
    _, cnt, ids = sim.read_spikes()
    rc, ri = ref.run_steps(steps, record=True)
    print(f"spikes gpu={cnt.sum()} oracle={rc.sum()}")
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))
    assert np.array_equal(sim.state(1), ref.state(1).astype(np.float32))
    assert np.array_equal(sim.syn_current(), ref.syn_current().astype(np.float32))
    assert np.array_equal(sim.pn_total(), ref.pn_total())
    return int(rc.sum())


@pytest.mark.parametrize("exec_mode", [1, 2], ids=["direct", "graph"])
def test_paired_cube_int_bit_exact(exec_mode):
    net = N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)
    assert delay1_per_wave(net).sum() > 0
    kw = dict(CFG0, int_weights=True, int_scale=65536.0)
    print("case: paired cube 10k", exec_mode, flush=True)
    assert check_int(net, 301, kw, exec_mode=exec_mode) > 500


@pytest.mark.parametrize("model", [0, 1], ids=["izh", "adex"])
def test_paired_grid_liquid_with_inputs_bit_exact(model):
    net, bits = grid_case(model)
    per_wave = delay1_per_wave(net)
    print(f"case: paired grid model {model}: delay-1 synapses per wave max {per_wave.max()}",
          flush=True)
    assert per_wave.max() > 64          # several pull rounds per wave
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=15.0, noise_kick_mv=20.0, noise_sd=0.5,
              seed=11, int_weights=True, int_scale=65536.0)
    assert check_int(net, 301, kw, series=bits, exp_mode=1) > 200


@pytest.mark.parametrize("scale", [4096.0, 65536.0], ids=["q12", "q16"])
def test_paired_all_delay1_bit_exact(scale):
    net = all_delay1_net(3000, 40, seed=5)
    print("case: paired all delay 1", scale, flush=True)
    kw = dict(CFG0, int_weights=True, int_scale=scale)
    assert check_int(net, 257, kw) > 500


def test_paired_float_ring_tolerance():
    """Float ring: v, u, I of every neuron whose spike count equals the
    oracle's within 1e-4 relative over 200 steps (denominator max(|x|, 1)),
    at most 0.1 % of neurons with a different count."""
    net = N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)
    kw = dict(CFG0, int_weights=False)
    sim = Sim(net, SimConfig(step_shape=2, **kw))
    assert sim.launch_shape() == 2
    ref = O.OracleSim(net, O.default_cfg(**kw))
    worst, diverged = 0.0, 0
    for _ in range(10):
        sim.run_steps(20)
        ref.run_steps(20)
        same = sim.pn_total() == ref.pn_total()
        diverged = max(diverged, int((~same).sum()))
        for g, r in ((sim.state(0), ref.state(0)), (sim.state(1), ref.state(1)),
                     (sim.syn_current(), ref.syn_current())):
            rel = np.abs(g.astype(np.float64) - r) / np.maximum(np.abs(r), 1.0)
            worst = max(worst, float(rel[same].max()))
    print(f"max rel error {worst:.2e}, diverged {diverged}")
    assert worst < 1e-4
    assert diverged <= 10


def test_paired_falls_back_when_watching():
    """A watched neuron needs the state after every step: the sim runs one
    step per launch while watching, paired again after, same raster."""
    net = N.cube_net(6000, mean_deg=80.0, seed=4, w_scale=0.7)
    kw = dict(CFG0, int_weights=True)
    sim = Sim(net, SimConfig(raster_steps=300, step_shape=2, **kw))
    assert sim.launch_shape() == 2
    sim.run_steps(100)
    sim.set_v_watch([3, 77, 5000])
    assert sim.launch_shape() == 1
    sim.run_steps(101)
    sim.set_v_watch([])
    assert sim.launch_shape() == 2
    sim.run_steps(99)
    _, cnt, ids = sim.read_spikes()
    rc, ri = O.OracleSim(net, O.default_cfg(**kw)).run_steps(300, record=True)
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)


def test_paired_batch_equals_alone():
    """Packed networks in paired launches: each network gets the raster of
    its one-step-per-launch run alone."""
    from neuro_edit_v2_amd.batch import Batch
    steps = 201
    nets, series = [], []
    for k, n in enumerate([700, 2197, 64, 1331, 3000]):
        nets.append(N.grid_net(0, model=k % 2, seed=k + 3, dt_ms=0.5, dims=W.grid_dims(n)))
        series.append(W.stream_window(W.rng_stream(100 + k, 8, 0.05), 8, steps))
    cfg = [dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=15.0 + k, noise_kick_mv=20.0,
                noise_sd=0.5, seed=11 * k + 1, int_weights=True, raster_steps=steps)
           for k in range(len(nets))]
    b = Batch(nets, [SimConfig(step_shape=2, **c) for c in cfg])
    for k in range(len(nets)):
        for q in range(8):
            b.play_input(k, q, series[k][:, q])
    b.run_steps(steps)
    for k, net in enumerate(nets):
        s = Sim(net, SimConfig(step_shape=1, **cfg[k]))
        for q in range(8):
            s.play_input(q, series[k][:, q])
        s.run_steps(steps)
        _, sc, si = s.read_spikes()
        s.close()
        _, bc, bi = b.read_spikes(k)
        assert np.array_equal(bc, sc) and np.array_equal(bi, si), f"net {k}"


def test_autotune_tries_paired():
    net = N.cube_net(5000, mean_deg=60.0, seed=21)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0, noise_sd=1.0, seed=3,
              int_weights=True)
    g = Sim(net, SimConfig(raster_steps=300, **kw))
    rep = g.autotune(64)
    print(rep)
    assert "paired" in rep and "chosen" in rep
    g.run_steps(300)
    _, c, ids = g.read_spikes()
    rc, ri = O.OracleSim(net, O.default_cfg(**kw)).run_steps(300, record=True)
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)


def test_paired_wait_gives_up_and_recomputes():
    """Give-up path of the pair wait, forced from the host: wave 0 withholds
    its published word and the poll limit is 64, so every wave it feeds
    gives up waiting and recomputes wave 0's first step of each launch from
    its launch-start state instead.  The run reports no error, the give-ups
    are counted, and the raster, v and I stay equal to the oracle's bit for
    bit; afterwards the ordinary path runs on, still exact."""
    net = all_delay1_net(3000, 40, seed=5)
    src0 = net["tgt"][: 64 * 40]                 # the synapses of wave 0's neurons
    assert len(np.unique(src0 // 64)) == (3000 + 63) // 64   # wave 0 feeds every wave
    kw = dict(CFG0, int_weights=True)
    sim = Sim(net, SimConfig(raster_steps=400, step_shape=2, **kw))
    assert sim.launch_shape() == 2
    n0 = C.c_uint64()
    L.check(L.lib().nh_test_pair_wait(sim._h, 64, 1, C.byref(n0)))
    print("case: paired wait, wave 0 muted, 64 polls", flush=True)
    sim.run_steps(20)
    sim.run_steps(180)
    n1 = C.c_uint64()
    L.check(L.lib().nh_test_pair_wait(sim._h, 0, 0, C.byref(n1)))   # synchronises
    assert n1.value > n0.value                   # the waits did give up
    ref = O.OracleSim(net, O.default_cfg(**kw))
    rc, ri = ref.run_steps(200, record=True)
    _, cnt, ids = sim.read_spikes()
    assert rc.sum() > 500
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))
    assert np.array_equal(sim.syn_current(), ref.syn_current().astype(np.float32))
    sim.run_steps(200)                           # ordinary waits again
    rc2, ri2 = ref.run_steps(200, record=True)
    _, cnt2, ids2 = sim.read_spikes()
    assert np.array_equal(cnt2, rc2) and np.array_equal(ids2, ri2)
    n2 = C.c_uint64()
    L.check(L.lib().nh_test_pair_wait(sim._h, 0, 0, C.byref(n2)))
    assert n2.value == n1.value


@pytest.mark.parametrize("window", [37, 64], ids=["odd_window", "even_window"])
def test_paired_with_raster_drains_mid_run(window):
    """A raster window smaller than the run makes the host drain it mid-run,
    which cuts chunks at arbitrary (also odd) lengths: pairs and one-step
    launches interleave, and the raster is still the oracle's."""
    net = N.cube_net(8000, mean_deg=80.0, seed=9, w_scale=0.7)
    kw = dict(CFG0, int_weights=True)
    sim = Sim(net, SimConfig(raster_steps=window, step_shape=2, **kw))
    assert sim.launch_shape() == 2
    for n in (101, 64, 136):
        sim.run_steps(n)
    _, cnt, ids = sim.read_spikes()
    rc, ri = O.OracleSim(net, O.default_cfg(**kw)).run_steps(301, record=True)
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert rc.sum() > 300


def random_net(seed):
    """A random network: size, out-degrees, delays (a quarter of them 1),
    weights, Izhikevich/AdEx mix all drawn from the seed."""
    r = np.random.default_rng(seed)
    n = int(r.integers(100, 7000))
    deg = r.poisson(r.uniform(5, 60), n).astype(np.int64)
    E = int(deg.sum())
    src_ptr = np.zeros(n + 1, np.uint32)
    src_ptr[1:] = np.cumsum(deg)
    maxd = int(r.integers(1, 32))
    delay = np.where(r.random(E) < 0.25, 1, r.integers(1, maxd + 1, E)).astype(np.uint16)
    w = (r.uniform(0.0, 3.0, E) * np.where(r.random(E) < 0.2, -1.0, 1.0)).astype(np.float32)
    adex = r.random() < 0.5
    if adex:
        model_ids = np.array([0, 1], np.uint8)
        model = (r.random(n) < 0.5).astype(np.uint8)
        P = 10
        p = np.zeros((n, P), np.float32)
        p[model == 0, :4] = [0.02, 0.2, -65.0, 8.0]
        p[model == 1] = [281.0, 30.0, -70.6, -50.4, 2.0, 4.0, 144.0, 80.5, -70.6, 20.0]
        x0 = np.zeros((n, 2), np.float32)
        x0[model == 0] = [-65.0, -13.0]
        x0[model == 1] = [-70.6, 0.0]
    else:
        model_ids = np.zeros(1, np.uint8)
        model = np.zeros(n, np.uint8)
        P = 4
        p = np.tile(np.float32([0.02, 0.2, -65.0, 8.0]), (n, 1))
        x0 = np.tile(np.float32([-65.0, -13.0]), (n, 1))
    return dict(n=n, src_ptr=src_ptr, tgt=r.integers(0, n, E, dtype=np.uint32), w=w, delay=delay,
                model_ids=model_ids, model=model, n_params=P, n_state=2,
                p=p.reshape(-1), x0=x0.reshape(-1))


@pytest.mark.parametrize("seed", [101, 202, 303, 404, 505, 606])
def test_paired_random_networks_bit_exact(seed):
    """Random networks (size, degrees, delays with many delay-1 synapses,
    weights, model mix, noise, ring scale, step count, workgroup size):
    the paired shape's raster and final state equal the oracle's."""
    r = np.random.default_rng(seed + 1)
    net = random_net(seed)
    kw = dict(dt_ms=float(r.choice([0.1, 0.25, 0.5])), synapse_tau_ms=float(r.uniform(1, 10)),
              noise_rate_hz=float(r.uniform(5, 60)), noise_kick_mv=20.0,
              noise_sd=float(r.choice([0.0, 0.5, 1.0])), seed=int(r.integers(1, 1 << 30)),
              int_weights=True, int_scale=float(r.choice([4096.0, 65536.0])))
    steps = int(r.integers(50, 400))
    block = int(r.choice([64, 128, 256, 512]))
    print(f"case: random net seed {seed}: n={net['n']} E={len(net['tgt'])} "
          f"d1={int((net['delay'] == 1).sum())} steps={steps} block={block} {kw}", flush=True)
    sim = Sim(net, SimConfig(raster_steps=steps, step_shape=2, block=block, **kw))
    assert sim.launch_shape() == 2
    sim.run_steps(steps)
    _, cnt, ids = sim.read_spikes()
    ref = O.OracleSim(net, O.default_cfg(exp_mode=1, **kw))
    rc, ri = ref.run_steps(steps, record=True)
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))
    assert np.array_equal(sim.state(1), ref.state(1).astype(np.float32))
    assert np.array_equal(sim.syn_current(), ref.syn_current().astype(np.float32))
