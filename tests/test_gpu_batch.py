"""Batched form: many uneven networks packed into one arena give each
network the raster it gets alone (int ring: bit-exact), against both the
plain HIP path and the oracle."""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import net as N
from neuro_edit_v2_amd import workloads as W
from neuro_edit_v2_amd.batch import Batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_batch(n_nets=10, steps=300, dt=0.5):
    sizes = [257, 1000, 64, 3000, 1331, 700, 4096, 123, 2197, 999][:n_nets]
    nets, series = [], []
    for k, n in enumerate(sizes):
        net = N.grid_net(0, model=k % 2, seed=k + 3, dt_ms=dt, dims=W.grid_dims(n))
        h = W.rng_stream(100 + k, 8, 0.05)
        bits = W.stream_window(h, 8, steps)
        nets.append(net)
        series.append(bits)
    return nets, series


def cfg_for(k, ring, steps):
    return SimConfig(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=15.0 + k, noise_kick_mv=20.0,
                     noise_sd=0.5, seed=11 * k + 1, int_weights=ring, raster_steps=steps)


@pytest.mark.parametrize("shape", [0, 1], ids=["two_kernel", "fused"])
def test_batch_equals_alone_int(shape):
    steps = 300
    nets, series = small_batch(steps=steps)
    cfgs = [cfg_for(k, True, steps) for k in range(len(nets))]
    for c in cfgs:
        c.step_shape = shape
    b = Batch(nets, cfgs)
    for k in range(len(nets)):
        for q in range(8):
            b.play_input(k, q, series[k][:, q])
    b.run_steps(steps)
    total = 0
    for k, net in enumerate(nets):
        _, bc, bi = b.read_spikes(k)
        s = Sim(net, cfgs[k])
        for q in range(8):
            s.play_input(q, series[k][:, q])
        s.run_steps(steps)
        _, sc, si = s.read_spikes()
        assert np.array_equal(bc, sc), f"net {k}: per-step counts differ"
        assert np.array_equal(bi, si), f"net {k}: ids differ"
        assert np.array_equal(b.state(k, 0), s.v())
        assert b.spikes_total(k) == s.spikes_total() == int(sc.sum())
        total += int(sc.sum())
        if k < 4:  # against the oracle too
            kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=15.0 + k, noise_kick_mv=20.0,
                      noise_sd=0.5, seed=11 * k + 1, int_weights=1, exp_mode=1)
            o = O.OracleSim(net, O.default_cfg(**kw))
            for q in range(8):
                o.play_input(q, series[k][:, q])
            oc, oi = o.run_steps(steps, record=True)
            assert np.array_equal(oc, bc) and np.array_equal(oi, bi)
    assert total > 100


def test_batch_float_ring_close_to_alone():
    steps = 200
    nets, series = small_batch(6, steps=steps)
    cfgs = [cfg_for(k, False, steps) for k in range(len(nets))]
    b = Batch(nets, cfgs)
    for k in range(len(nets)):
        for q in range(8):
            b.play_input(k, q, series[k][:, q])
    b.run_steps(steps)
    for k, net in enumerate(nets):
        s = Sim(net, cfgs[k])
        for q in range(8):
            s.play_input(q, series[k][:, q])
        s.run_steps(steps)
        a, c = b.spikes_total(k), s.spikes_total()
        assert abs(a - c) <= max(3, 0.05 * c), (k, a, c)


def test_batch_autotune_keeps_rasters_bit_exact():
    """The batch autotuner times launch shapes, workgroup sizes and graph
    lengths from step 0 and resets (stimulus programs included); whatever
    it picks, every network keeps the raster it gets alone."""
    steps = 300
    nets, series = small_batch(steps=steps)
    cfgs = [cfg_for(k, True, steps) for k in range(len(nets))]
    b = Batch(nets, cfgs)
    for k in range(len(nets)):
        for q in range(8):
            b.play_input(k, q, series[k][:, q])
    rep = b.autotune(64)
    print(rep)
    assert "chosen" in rep
    b.run_steps(steps)
    for k, net in enumerate(nets):
        _, bc, bi = b.read_spikes(k)
        s = Sim(net, cfgs[k])
        for q in range(8):
            s.play_input(q, series[k][:, q])
        s.run_steps(steps)
        _, sc, si = s.read_spikes()
        s.close()
        assert np.array_equal(bc, sc) and np.array_equal(bi, si), f"net {k}"
