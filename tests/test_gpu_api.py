"""GPU: the Backend-trait surface of the HIP path against the oracle, on the
reference's own known-answer scenarios (sim/tests/cpu_ref.rs,
spike_ring.rs, adex.rs, v_watch.rs, grid_trace.rs, robustness.rs).

Every comparison runs the int ring or a spike-free path where the float
arithmetic is identical, so the bar is bit equality unless stated."""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import net as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = pytest.mark.parametrize("shape", [0, 1], ids=["two_kernel", "fused"])


def line(n):
    b = O.NetBuilder()
    for i in range(n):
        b.add_neuron((i * 50.0, 0.0, 0.0), O.NetBuilder.EXC)
    return b


def pair_sims(net, shape=1, **kw):
    g = Sim(net, SimConfig(step_shape=shape, **kw))
    o = O.OracleSim(net, O.default_cfg(**{k: v for k, v in kw.items()
                                          if k not in ("raster_steps", "exec_mode", "graph_steps")}))
    return g, o


def f32eq(g, o):
    return np.array_equal(np.asarray(g, np.float32), np.asarray(o, np.float64).astype(np.float32))


@SHAPES
def test_hand_computed_izh_steps(shape):
    """cpu_ref.rs:17-46 on the GPU: v = -64.3, u = -12.99972 after one step."""
    g, o = pair_sims(line(1).build(0.1), shape)
    for s in (g, o):
        s.set_external(0, 10.0)
        s.run_steps(1)
    assert abs(g.v()[0] - (-64.3)) < 1e-4 and abs(g.state(1)[0] - (-12.99972)) < 1e-4
    assert f32eq(g.v(), o.v()) and f32eq(g.state(1), o.state(1))
    for s in (g, o):
        s.run_steps(20)
    assert f32eq(g.v(), o.v())


@SHAPES
@pytest.mark.parametrize("ring", [False, True], ids=["float", "int"])
def test_delay_two_arrival_and_decay(shape, ring):
    """cpu_ref.rs:72-114: arrival after exactly 2 steps, s = 30 e^-0.05."""
    b = line(2)
    b.push_synapse(0, 1, 30.0, 0.2)
    g, o = pair_sims(b.build(0.1), shape, int_weights=ring)
    for s in (g, o):
        s.set_external(0, 25.0)
    spike_t = None
    for _ in range(500):
        before = o.spikes_total()
        g.run_steps(1)
        o.run_steps(1)
        assert g.spikes_total() == o.spikes_total()
        if o.spikes_total() > before:
            spike_t = o.t - 1
            break
    assert spike_t is not None
    for s in (g, o):
        s.set_external(0, 0.0)
        s.run_steps(1)
    assert g.syn_current()[1] == 0.0
    for s in (g, o):
        s.run_steps(1)
    assert abs(g.syn_current()[1] - 30.0 * np.exp(-0.05)) < 1e-3
    assert f32eq(g.syn_current(), o.syn_current())
    for s in (g, o):
        s.run_steps(5)
    assert f32eq(g.syn_current(), o.syn_current()) and f32eq(g.v(), o.v())


def test_stimulus_constant_and_played_series():
    """cpu_ref.rs:258-303: routed input constants and step-locked series."""
    b = line(3)
    b.add_input([(0, 1.0), (2, 2.0)])
    b.add_input([(1, 1.0)])
    g, o = pair_sims(b.build(0.1))
    for s in (g, o):
        s.set_input_constant(0, 5.0)
        s.run_steps(1)
    assert abs(g.v()[0] - (-64.8)) < 0.2 and f32eq(g.v(), o.v())
    assert np.array_equal(g.input_values(), [5.0, 0.0])
    for s in (g, o):
        s.play_input(1, [10.0, 0.0, 0.0, 25.0, 25.0])
        s.loop_input(0, [1.0, -2.0, 3.0])
        s.run_steps(40)
    assert f32eq(g.v(), o.v()) and f32eq(g.state(1), o.state(1))


def test_stimulus_chunks_longer_than_window():
    """A played series spanning several host upload windows (1024 steps)."""
    b = line(2)
    b.add_input([(0, 1.0), (1, 0.5)])
    g, o = pair_sims(b.build(0.1))
    ser = (np.sin(np.arange(3000) * 0.01) * 20.0).astype(np.float32)
    for s in (g, o):
        s.play_input(0, ser)
        s.run_steps(2500)
    assert f32eq(g.v(), o.v())
    assert g.spikes_total() == o.spikes_total() > 0


def test_reset_clears_state_and_replays_identically():
    """cpu_ref.rs:323-334 plus: a reset run repeats bit for bit."""
    net = N.cube_net(2000, mean_deg=50.0, seed=9)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0, seed=3, int_weights=True,
              raster_steps=300)
    g = Sim(net, SimConfig(**kw))
    g.run_steps(300)
    _, c1, i1 = g.read_spikes()
    v1 = g.v()
    g.reset_state()
    assert g.step_count() == 0 and np.all(g.v()[:5] == -65.0) and g.spikes_total() == 0
    assert np.all(g.trace() == 0.0)
    g.run_steps(300)
    _, c2, i2 = g.read_spikes()
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2) and np.array_equal(v1, g.v())


@SHAPES
def test_trace_readout_bit_exact(shape):
    """cpu.rs:252-258 trace rebuilt from last-spike steps equals the
    oracle's running trace bit for bit (int ring, noisy net)."""
    net = N.cube_net(3000, mean_deg=50.0, seed=2)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, readout_tau_ms=7.0, noise_rate_hz=40.0, seed=5,
              int_weights=True)
    g, o = pair_sims(net, shape, **kw)
    for n in (1, 37, 250):
        g.run_steps(n)
        o.run_steps(n)
        assert f32eq(g.trace(), o.trace())


def test_spike_ring_wrap_and_window_semantics():
    """spike_ring.rs:25-59 on the GPU (cuda variants :136-166)."""
    net = line(1).build(0.1)
    g = Sim(net, SimConfig(record_spikes=True, spike_record_steps=64))
    g.set_external(0, 40.0)
    g.run_steps(64 * 20)
    g.set_external(0, 0.0)
    g.run_steps(200)
    before = g.spikes_total()
    g.run_steps(64)
    assert g.spike_counts(64)[0] == g.spikes_total() - before
    g2 = Sim(net, SimConfig(record_spikes=True, spike_record_steps=64))
    g2.set_external(0, 40.0)
    g2.run_steps(300)
    direct = int(g2.spike_counts(5000)[0])
    assert direct == g2.spikes_total() > 0
    assert int(g2.spike_counts(64)[0]) <= direct
    with pytest.raises(ValueError, match="exceeds"):
        g2.spike_counts(200)


def test_spike_rows_and_counts_match_oracle():
    net = N.cube_net(1500, mean_deg=40.0, seed=4)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=60.0, seed=8, int_weights=True,
              record_spikes=True, spike_record_steps=100)
    g, o = pair_sims(net, **kw)
    g.run_steps(333)
    o.run_steps(333)
    for back in (0, 1, 50, 99):
        assert np.array_equal(g.spike_row(back), o.spike_row(back))
    assert g.spike_row(100) is None and o.spike_row(100) is None
    assert np.array_equal(g.spike_counts(80), o.spike_counts(80))
    assert np.array_equal(g.spike_counts(10_000), o.spike_counts(10_000))
    row = o.spike_row(0)
    ids = [wi * 64 + b for wi, w in enumerate(row) for b in range(64) if (int(w) >> b) & 1]
    assert g.last_spikes().tolist() == ids
    assert abs(g.mean_rate_hz() - o.spikes_total() / 1500 / (333 * 0.5 / 1000)) < 1e-9


def test_grid_traces_bit_exact():
    """cpu.rs:361-407: per-texel exp(-(t - last)*dt/tau) with libm expf."""
    net = N.cube_net(1000, mean_deg=30.0, seed=6)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=50.0, seed=2, int_weights=True)
    g, o = pair_sims(net, **kw)
    g.run_steps(50)
    o.run_steps(50)
    groups = [list(range(0, 300)), list(range(500, 700)) + [3, 3, 999]]
    for s in (g, o):
        s.set_grid_watches(groups)
        s.run_steps(120)
    for tau in (3.0, 25.0):
        for a, b in zip(g.grid_traces(tau), o.grid_traces(tau)):
            assert np.array_equal(a, b)


def test_v_watch_drain_is_per_step_state():
    """v_watch.rs:43-70: drained samples are the per-step state rows."""
    net = N.cube_net(800, mean_deg=30.0, seed=3)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=80.0, seed=4, int_weights=True)
    g, o = pair_sims(net, **kw)
    ids = [0, 17, 799]
    g.set_v_watch(ids)
    want_v, want_u = [], []
    for _ in range(25):
        o.run_steps(1)
        want_v.append(o.v()[ids])
        want_u.append(o.state(1)[ids])
    g.run_steps(25)
    dv, du = g.drain_watch()
    for c in range(3):
        assert f32eq(dv[c], np.array(want_v)[:, c]) and f32eq(du[c], np.array(want_u)[:, c])
    assert all(len(x) == 0 for x in g.drain_watch()[0])


def test_set_noise_live_and_gaussian_sd():
    net = N.cube_net(2000, mean_deg=40.0, seed=1)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, seed=2, int_weights=True, noise_sd=2.0)
    g, o = pair_sims(net, **kw)
    for s in (g, o):
        s.run_steps(50)
        s.set_noise(25.0, 15.0)
        s.run_steps(100)
    assert f32eq(g.v(), o.v()) and g.spikes_total() == o.spikes_total() > 0


@pytest.mark.parametrize("exp_mode_ok", [True])
def test_adex_mixed_net_bit_exact(exp_mode_ok):
    """adex.rs:56-121: mixed izh/AdEx net; identical spike counts (here:
    identical raster and state against the oracle's portable-exp mode)."""
    b = line(8)
    tonic = [281.0, 30.0, -70.6, -50.4, 2.0, 4.0, 144.0, 80.5, -70.6, 20.0]
    for i in (1, 3, 4, 6):
        b.set_model(i, 1, tonic)
    b.connect(0, 1, 8.0)
    b.connect(1, 2, 8.0)
    b.connect(3, 4, 5.0)
    net = b.build(0.1)
    for ring in (False, True):
        g = Sim(net, SimConfig(record_spikes=True, int_weights=ring))
        o = O.OracleSim(net, O.default_cfg(record_spikes=1, int_weights=int(ring), exp_mode=1))
        for i, a in [(0, 15.0), (1, 900.0), (3, 800.0), (4, 1000.0), (6, 1200.0)]:
            g.set_external(i, a)
            o.set_external(i, a)
        g.run_steps(3000)
        o.run_steps(3000)
        gc, oc = g.spike_counts(1 << 62), o.spike_counts(1 << 62)
        assert np.array_equal(gc, oc)
        assert all(gc[i] >= 2 for i in (1, 3, 4, 6))
        assert f32eq(g.v(), o.v())


def test_adex_grid_net_int_raster():
    net = N.grid_net(12, model=1, seed=7, dt_ms=0.5)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=20.0, seed=4, int_weights=True)
    g = Sim(net, SimConfig(raster_steps=400, **kw))
    o = O.OracleSim(net, O.default_cfg(exp_mode=1, **kw))
    for k in range(8):
        bits = (np.random.default_rng(k).random(400) < 0.05).astype(np.float32)
        g.play_input(k, bits)
        o.play_input(k, bits)
    g.run_steps(400)
    _, c, ids = g.read_spikes()
    rc, ri = o.run_steps(400, record=True)
    assert rc.sum() > 0
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)


def test_per_neuron_param_rows():
    """More than 256 distinct rows: per-neuron parameter mode."""
    net = N.cube_net(3000, mean_deg=30.0, seed=11)
    p = net["p"].reshape(-1, 4).copy()
    p[:, 3] = np.linspace(2.0, 8.0, len(p)).astype(np.float32)
    net["p"] = p.ravel()
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=40.0, seed=1, int_weights=True)
    g, o = pair_sims(net, raster_steps=200, **kw)
    g.run_steps(200)
    _, c, ids = g.read_spikes()
    rc, ri = o.run_steps(200, record=True)
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)


def test_table_params_more_than_kernel_argument_rows():
    """5..256 distinct rows: the global parameter table."""
    net = N.cube_net(3000, mean_deg=30.0, seed=12)
    p = net["p"].reshape(-1, 4).copy()
    p[:, 2] = (-65.0 + (np.arange(len(p)) % 7)).astype(np.float32)
    net["p"] = p.ravel()
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=40.0, seed=1, int_weights=True)
    g, o = pair_sims(net, raster_steps=200, **kw)
    g.run_steps(200)
    _, c, ids = g.read_spikes()
    rc, ri = o.run_steps(200, record=True)
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)


def test_long_delays_unpacked_layout():
    """Delays beyond 31 steps use the unpacked synapse layout."""
    net = N.bench_net(2000, 50, 3, 0.7, 5.0, 0.1)   # delays up to 210 steps
    kw = dict(dt_ms=0.1, noise_rate_hz=300.0, seed=2, int_weights=True)
    g, o = pair_sims(net, raster_steps=700, **kw)
    g.run_steps(700)
    _, c, ids = g.read_spikes()
    rc, ri = o.run_steps(700, record=True)
    assert rc.sum() > 100
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)


def test_empty_net_builds_and_runs():
    """robustness.rs:8-40: an empty net must build and run."""
    net = dict(n=0, src_ptr=np.zeros(1, np.uint32), tgt=np.zeros(0, np.uint32),
               w=np.zeros(0, np.float32), delay=np.zeros(0, np.uint16),
               model_ids=np.zeros(1, np.uint8), model=np.zeros(0, np.uint8), n_params=4,
               n_state=2, p=np.zeros(0, np.float32), x0=np.zeros(0, np.float32))
    g = Sim(net, SimConfig())
    g.run_ms(10.0)
    assert g.step_count() == 100 and g.spikes_total() == 0 and len(g.v()) == 0


def test_dev_mem_and_time():
    net = N.cube_net(1000, mean_deg=20.0, seed=1)
    g = Sim(net, SimConfig(dt_ms=0.5))
    assert g.dev_mem_bytes() > 1000 * 20 * 8
    g.run_ms(7.0)
    assert g.step_count() == 14 and abs(g.time_ms() - 7.0) < 1e-9


def test_autotune_keeps_results_bit_exact():
    """The autotuner times launch shapes and workgroup sizes from step 0 and
    resets; whatever it picks, the raster stays the oracle's."""
    net = N.cube_net(5000, mean_deg=60.0, seed=21)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0, noise_sd=1.0, seed=3,
              int_weights=True)
    g = Sim(net, SimConfig(raster_steps=300, **kw))
    rep = g.autotune(64)
    assert "chosen" in rep and g.step_count() == 0
    g.run_steps(300)
    _, c, ids = g.read_spikes()
    o = O.OracleSim(net, O.default_cfg(**kw))
    rc, ri = o.run_steps(300, record=True)
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)


@pytest.mark.parametrize("shape", [0, 1], ids=["two_kernel", "fused"])
@pytest.mark.parametrize("block", [64, 128, 512])
def test_workgroup_sizes_bit_exact(block, shape):
    net = N.cube_net(4000, mean_deg=60.0, seed=22)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0, seed=5, int_weights=True)
    g = Sim(net, SimConfig(raster_steps=200, block=block, step_shape=shape, **kw))
    g.run_steps(200)
    _, c, ids = g.read_spikes()
    rc, ri = O.OracleSim(net, O.default_cfg(**kw)).run_steps(200, record=True)
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)


def lsm_like(n_side, seed):
    """sim/tests/gpu_vs_cpu.rs:11-18"""
    b = O.NetBuilder()
    rng = O.Rng(seed)
    g = b.grid3d(n_side, n_side, n_side, 60.0)
    b.set_exc_inh_exact(g, n_side ** 3 // 5, rng)
    b.create_connections(g, 0.4, 100.0, 5.0, rng)
    return b.build(0.1)


@SHAPES
@pytest.mark.parametrize("ring", [False, True], ids=["float", "int"])
def test_custom_tau_single_synapse(shape, ring):
    """cpu_ref.rs:117-206: a tau=20 synapse feeds its own trace; the GPU
    state and spikes follow the oracle bit for bit."""
    b = line(2)
    b.push_synapse(0, 1, 30.0, 0.2, 20.0)
    g, o = pair_sims(b.build(0.1), shape, int_weights=ring)
    for s in (g, o):
        s.set_external(0, 25.0)
        s.run_steps(40)
        s.set_external(0, 0.0)
        s.run_steps(200)
    assert f32eq(g.v(), o.v()) and f32eq(g.state(1), o.state(1))
    assert g.spikes_total() == o.spikes_total()


def test_custom_tau_net_matches_oracle():
    """gpu_vs_cpu.rs:51-80: per-synapse tau on the recurrent edges
    (Gaussian 25 +- 10 ms, clamped at 0 = merged path); int ring raster and
    state bit-exact."""
    net = lsm_like(8, 42)
    r = np.random.default_rng(9)
    net["tau"] = np.maximum(r.normal(25.0, 10.0, len(net["tgt"])), 0.0).astype(np.float32)
    net["tau"][::7] = 0.0
    kw = dict(record_spikes=True, int_weights=True, noise_rate_hz=300.0, seed=4)
    g = Sim(net, SimConfig(raster_steps=500, **kw))
    o = O.OracleSim(net, O.default_cfg(**kw))
    for i in range(8):
        g.set_external(i, 12.0)
        o.set_external(i, 12.0)
    g.run_steps(500)
    _, c, ids = g.read_spikes()
    rc, ri = o.run_steps(500, record=True)
    assert rc.sum() > 0
    assert np.array_equal(c, rc) and np.array_equal(ids, ri)
    assert f32eq(g.v(), o.v())
    assert g.events_total() == o.events_total()


@pytest.mark.parametrize("shape", [1, 2], ids=["fused", "paired"])
def test_counter_mark_counts_the_window_exactly(shape):
    """nh_mark_counters / nh_counts_since_mark (the benchmark's window
    counters, a device-side snapshot): spikes and events since the mark equal
    the difference of the totals read on the host; a reset clears the mark."""
    from neuro_edit_v2_amd import _lib as L
    net = N.cube_net(6000, mean_deg=80.0, seed=4, w_scale=0.7)
    sim = Sim(net, SimConfig(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0, noise_kick_mv=20.0,
                             seed=3, int_weights=True, step_shape=shape))
    sim.run_steps(51)
    s0, e0 = sim.spikes_total(), sim.events_total()
    sim.mark_counters()
    sim.run_steps(77)
    sp, ev = sim.counts_since_mark()
    assert sp > 0 and sp == sim.spikes_total() - s0
    assert ev == sim.events_total() - e0
    sim.reset_state()
    with pytest.raises(L.NHError):
        sim.counts_since_mark()
    sim.mark_counters()
    assert sim.counts_since_mark() == (0, 0)


@pytest.mark.parametrize("shape", [1, 2], ids=["fused", "paired"])
def test_spike_counts_counted_on_device_match_oracle(shape):
    """Backend::spike_counts (cuda.rs:725-763; cpu.rs:452-486) with the
    window counted on the device: rows within the record ring (every
    remainder of the kernel's eight-row unroll, windows that wrap the
    ring), rows >= t (lifetime counts) and the error between the two."""
    net = N.cube_net(5000, mean_deg=60.0, seed=12)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=40.0, seed=3, int_weights=True,
              record_spikes=True, spike_record_steps=100)
    g, o = pair_sims(net, shape=shape, **kw)
    for steps in (5, 233, 120):
        g.run_steps(steps)
        o.run_steps(steps)
        t = o.t
        for rows in (1, 2, 7, 8, 9, 15, 16, 17, 63, 64, 99, 100, t, t + 1, 1 << 40):
            if rows > 100 and rows < t:
                continue
            assert np.array_equal(g.spike_counts(rows), o.spike_counts(rows)), (t, rows)
        for rows in (101, t - 1):
            if 100 < rows < t:
                with pytest.raises(ValueError, match="exceeds"):
                    g.spike_counts(rows)
    assert int(o.spike_counts(100).sum()) > 100


def test_batch_spike_counts_per_network_match_oracle():
    """nh_batch_spike_counts: each packed network's window counts equal the
    oracle's for that network alone (int ring)."""
    from neuro_edit_v2_amd import workloads as W
    from neuro_edit_v2_amd.batch import Batch
    sizes = [257, 1000, 64, 3000, 700]
    steps = 160
    nets, series, cfgs = [], [], []
    for k, n in enumerate(sizes):
        nets.append(N.grid_net(0, model=k % 2, seed=k + 3, dt_ms=0.5, dims=W.grid_dims(n)))
        series.append(W.stream_window(W.rng_stream(100 + k, 8, 0.05), 8, steps))
        cfgs.append(SimConfig(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0, seed=k + 1,
                              int_weights=True, record_spikes=True, spike_record_steps=64))
    b = Batch(nets, cfgs)
    for k in range(len(nets)):
        for q in range(8):
            b.play_input(k, q, series[k][:, q])
    b.run_steps(steps)
    for k, net in enumerate(nets):
        o = O.OracleSim(net, O.default_cfg(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0,
                                           seed=k + 1, int_weights=1, exp_mode=1,
                                           record_spikes=1, spike_record_steps=64))
        for q in range(8):
            o.play_input(q, series[k][:, q])
        o.run_steps(steps)
        for rows in (1, 13, 64, steps, 1 << 40):
            assert np.array_equal(b.spike_counts(k, rows), o.spike_counts(rows)), (k, rows)
        with pytest.raises(ValueError, match="exceeds"):
            b.spike_counts(k, 65)
